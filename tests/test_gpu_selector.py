"""Learned format selector on the GPU: structure features against the
oracle's definitions, the HIP backward pass against the oracle's gradient,
one Adam step, BF16 inference on the matrix cores, and a held-out regret test
of a network trained on timings measured here."""
import json
import os

import numpy as np
import pytest

import oracle
from gpu_cases import announce, case
from oracle import selector as S

pytestmark = pytest.mark.gpu

# MFP_RESULTS_DIR=<dir>: also write the evaluation summary there
OUT = os.environ.get("MFP_RESULTS_DIR")


@pytest.mark.parametrize("name", ["stencil2d_64", "stencil3d_20", "powerlaw_20k", "uniform_band", "edge_rows",
                                  "one_by_one"])
def test_features_match_oracle(gpu, name):
    rp, ci, va, cols = case(name, "float32")
    announce(f"features case={name}")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    f = gpu.selector.features(A)
    ref = S.features(rp, ci, rp.size - 1, cols)
    # the device counters are exact; the last rounding to float32 may differ
    assert np.allclose(f, ref.astype(np.float32), rtol=2e-7, atol=1e-9), np.nonzero(
        ~np.isclose(f, ref, rtol=2e-7, atol=1e-9))


def _problem(seed, n):
    rng = np.random.default_rng(seed)
    X = rng.uniform(0, 1, (n, 64)).astype(np.float32)
    P = rng.dirichlet(np.ones(8) * 0.3, n).astype(np.float32)
    return X, P


@pytest.mark.parametrize("n", [1, 37, 256])
def test_gradient_matches_oracle(gpu, n):
    announce(f"mlp gradient n={n}")
    net = gpu.selector.MLP(seed=n)
    theta = net.params
    X, P = _problem(n, n)
    loss, g = net.loss_grad(X, P)
    ref_loss, ref_g = S.loss_grad(theta.astype(np.float64), X, P)
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    # FP32 forward/backward against float64: relative L2 per parameter block
    o = 0
    for name, shape in S.SHAPES:
        k = int(np.prod(shape))
        a, b = g[o:o + k].astype(np.float64), ref_g[o:o + k]
        assert np.linalg.norm(a - b) <= 1e-4 * max(np.linalg.norm(b), 1e-12), name
        o += k
    net.destroy()


def test_adam_steps_match_oracle(gpu):
    announce("mlp adam")
    net = gpu.selector.MLP(seed=5)
    theta = net.params.astype(np.float64)
    X, P = _problem(9, 64)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    for step in (1, 2, 3):
        _, g = S.loss_grad(theta, X, P)
        theta, m, v = S.adam(theta, g, m, v, step, 1e-3)
        net.train_step(X, P, 1e-3)
    got = net.params.astype(np.float64)
    assert np.max(np.abs(got - theta)) < 1e-5
    net.destroy()


def test_set_params_round_trip_and_reset(gpu):
    announce("mlp params")
    net = gpu.selector.MLP(seed=1)
    theta = np.random.default_rng(0).normal(0, 0.1, net.nparam).astype(np.float32)
    net.params = theta
    assert np.array_equal(net.params, theta)
    with pytest.raises(ValueError):
        net.params = theta[:-1]
    net.destroy()


@pytest.mark.parametrize("n", [1, 16, 100])
def test_bf16_inference_matches_emulation(gpu, n):
    announce(f"mlp bf16 inference n={n}")
    net = gpu.selector.MLP(seed=2)
    theta = net.params
    X, _ = _problem(100 + n, n)
    q = net.predict(X, bf16=True)
    ref = S.predict_bf16(theta, X)
    assert np.allclose(q.sum(axis=1), 1.0, atol=1e-6)
    assert np.max(np.abs(q - ref)) < 2e-3
    q32 = net.predict(X)
    _, _, Z = S.forward(theta.astype(np.float64), X)
    assert np.max(np.abs(q32 - S.softmax(Z))) < 1e-5
    net.destroy()


def test_autotune_marks_each_formats_default(gpu):
    rp, ci, va, cols = case("uniform_band", "float32")
    announce("autotune defaults")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    x = gpu.DeviceArray.from_host(gpu.generate_vector(1, cols, np.float32))
    y = gpu.DeviceArray(rp.size - 1, np.float32)
    res = gpu.autotune(A, x, y, reps=1)
    dflt = [r for r in res if r["tuning"]["flags"] & gpu._lib.MFP_TUNE_IS_DEFAULT]
    assert sorted(r["format"] for r in dflt) == sorted(gpu.FORMATS)
    for r in dflt:
        t = gpu.engine.default_tuning(A, r["format"])
        for k in ("lanes", "nnz_per_wg", "hyb_width", "scoo_slice_size", "scoo_lanes", "scoo_slices"):
            assert r["tuning"][k] == getattr(t, k), (r["format"], k)


def test_selector_held_out_regret(gpu, tmp_path):
    """Sweep 120 small matrices, train on 90 as the shipped selector is
    trained (tuning rules fitted first, the network on the rules' targets),
    evaluate on 30 held out: the pick (format by the network, tuning by the
    rules) against the exhaustive best, next to the best single fixed format
    chosen on the training set, with default tuning and with the same rules.
    The network with default tuning throughout is reported beside them."""
    sel = gpu.selector
    path = str(tmp_path / "sweep.jsonl")
    announce("selector sweep 120 matrices")
    n_meas = sel.sweep(path, 120, seed=11, max_nnz=1 << 20, reps=3, log=None)
    assert n_meas == 120
    assert sel.sweep(path, 120, seed=11, max_nnz=1 << 20, log=None) == 0  # resumes: nothing left
    recs = sel.load_records(path)
    assert [r["index"] for r in recs] == list(range(120))
    # the sampler keeps to its bounds
    for r in recs:
        assert r["nnz"] <= 2 * (1 << 20) and r["rows"] >= 2000
    tr, te = recs[:90], recs[90:]
    rules = sel.fit_tuning_rules(tr)
    X, P = sel.dataset(tr, rules)
    announce("selector train")
    net = sel.train(X, P, epochs=400, batch=32, lr=2e-3, seed=0)
    Xt, _ = sel.dataset(te, rules)
    rep_rules32 = sel.evaluate_with_rules(te, net.predict(Xt), rules)
    rep_rules = sel.evaluate_with_rules(te, net.predict(Xt, bf16=True), rules)
    rep_default = sel.evaluate(te, net.predict(Xt, bf16=True))
    fixed_default = sel.fixed_format_baseline(tr, te)
    fixed_rules = sel.fixed_format_baseline(tr, te, rules)
    fixed_regret = min(fixed_default["mean_regret"], fixed_rules["mean_regret"])
    summary = {"held_out_with_rules": rep_rules32, "held_out_bf16_with_rules": rep_rules,
               "held_out_bf16_default_tuning": rep_default,
               "fixed_default": fixed_default, "fixed_rules": fixed_rules,
               "autotune_s_mean": float(np.mean([r["autotune_s"] for r in recs])),
               "feature_s_mean": float(np.mean([r["feature_s"] for r in recs]))}
    if OUT:
        os.makedirs(OUT, exist_ok=True)
        with open(os.path.join(OUT, "selector_small_eval.json"), "w") as f:
            json.dump(summary, f, indent=1)
    print(json.dumps(summary))
    assert np.isfinite(rep_rules["mean_regret"])
    # what the selector gains over the method without it (one fixed format,
    # with default tuning or with the same tuning rules)
    assert rep_rules32["mean_regret"] < fixed_regret
    assert rep_rules["mean_regret"] < fixed_regret
    # the saved selector picks a (format, tuning) that builds and multiplies correctly
    path_sel = str(tmp_path / "sel.npz")
    sel.Selector(net, rules).save(path_sel)
    loaded = sel.Selector.load(path_sel)
    assert np.array_equal(loaded.net.params, net.params)
    rp, ci, va, cols = case("powerlaw_20k", "float32")
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, cols)
    fmt, tuning = loaded.pick(A)
    assert fmt in sel.FORMAT_ORDER
    x = gpu.generate_vector(2, cols, np.float32)
    xd, yd = gpu.DeviceArray.from_host(x), gpu.DeviceArray(rp.size - 1, np.float32)
    gpu.Plan(A, fmt, **tuning).spmv(xd, yd)
    gpu.synchronize()
    y64, absdot = oracle.spmv_f64(rp, ci, va, x)
    assert oracle.tolerance_check(yd.to_host(), y64, absdot, 1e-5) <= 1e-5
    net.destroy()
