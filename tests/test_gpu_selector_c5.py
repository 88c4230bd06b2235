"""The shipped learned selector at the selector configuration's scale.

The committed sweep (results/selector_c5_sweep*.jsonl.gz: 2,400 matrices,
exhaustive timings of every format and tuning on an MI355X) is split as the
selector was trained: every fifth record held out (480), the rest (1,920)
trained matrix_format_performance_amd/selector_mi355x.npz.  Here:

- the network's BF16 inference runs on the matrix cores for the 480 held-out
  feature vectors; the picks (format + the tuning rules' tuning) must be
  within 5% of the exhaustive best on average, and beat the method without a
  selector (the one fixed format best on the training matrices);
- every 8th held-out matrix (60) is generated again from its recorded
  spec and seed: its size and its 64 features, computed on the GPU from
  integer counters, must equal the record exactly, and a fresh exhaustive
  timing of every candidate must keep the pick within 5% of the best on
  average.
"""
import gzip
import json
import math
import os

import numpy as np
import pytest

from gpu_cases import announce

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SWEEPS = [os.path.join(ROOT, "results", n) for n in ("selector_c5_sweep.jsonl.gz", "selector_c5_sweep_big.jsonl.gz")]
HOLD_EVERY = 5
RETIME_EVERY = 8


def _load(path):
    recs = {}
    with gzip.open(path, "rt") as f:
        for line in f:
            r = json.loads(line)
            recs[r["index"]] = r
    return [recs[k] for k in sorted(recs)]


@pytest.fixture(scope="module")
def split():
    recs = [r for p in SWEEPS for r in _load(p)]
    assert len(recs) == 2400
    tr = [r for i, r in enumerate(recs) if i % HOLD_EVERY != HOLD_EVERY - 1]
    te = [r for i, r in enumerate(recs) if i % HOLD_EVERY == HOLD_EVERY - 1]
    return tr, te


@pytest.fixture(scope="module")
def shipped(gpu):
    return gpu.selector.Selector.default()


def _pick_times(sel_mod, records, probs, rules):
    out = []
    for r, j in zip(records, np.argmax(probs, axis=1)):
        f = sel_mod.FORMAT_ORDER[j]
        avail = {c[1] for c in r["cands"] if c[0] == f}
        out.append(sel_mod.record_time(r, f, sel_mod.choose_label(rules, f, r["features"], avail))
                   if avail else math.inf)
    return np.array(out)


def test_held_out_regret_with_gpu_bf16_inference(gpu, split, shipped):
    S = gpu.selector
    tr, te = split
    announce(f"selector C5: BF16 inference on {len(te)} held-out records")
    X = np.array([r["features"] for r in te], np.float32)
    probs = shipped.net.predict(X, bf16=True)
    ev = S.evaluate_with_rules(te, probs, shipped.rules)
    base_rules = S.fixed_format_baseline(tr, te, shipped.rules)
    base_default = S.fixed_format_baseline(tr, te)
    print(json.dumps({"held_out": ev, "fixed_rules": base_rules, "fixed_default": base_default}))
    assert ev["mean_regret"] <= 0.05
    assert ev["mean_regret"] < base_rules["mean_regret"] and ev["mean_regret"] < base_default["mean_regret"]
    # FP32 inference picks the same formats on all but a few near ties
    p32 = shipped.net.predict(X, bf16=False)
    assert np.mean(np.argmax(p32, axis=1) == np.argmax(probs, axis=1)) >= 0.98


RETIMED = {}


def _retime_indices():
    return list(range(0, 480, RETIME_EVERY))


@pytest.mark.parametrize("k", _retime_indices())
def test_regenerated_matrix_features_and_fresh_regret(gpu, split, shipped, k):
    S = gpu.selector
    _, te = split
    r = te[k]
    sp = r["spec"]
    spec = gpu.GenSpec(r["family"], values=["uniform", "normal"][r["index"] % 2], seed=r["index"], n=sp["n"],
                       rows=sp["rows"], nnz_lo=sp["nnz_lo"], nnz_hi=sp["nnz_hi"], alpha=sp["alpha"],
                       bandwidth=sp["bandwidth"])
    announce(f"selector C5: held-out {k} (sweep index {r['index']}, {r['family']}, rows {r['rows']}, nnz {r['nnz']})")
    rp, ci, va = gpu.generate(spec, np.float32)
    assert (rp.size - 1, ci.size) == (r["rows"], r["nnz"])
    A = gpu.DeviceMatrix.from_csr(rp, ci, va, r["cols"])
    f = S.features(A)
    assert np.array_equal(f, np.array(r["features"], np.float32))
    fmt, tuning = shipped.pick(A, row_ptr=rp)
    x = gpu.DeviceArray.from_host(gpu.generate_vector(1, r["cols"], np.float32))
    y = gpu.DeviceArray(r["rows"], np.float32)
    cands = gpu.autotune(A, x, y, reps=3, flush=True)
    ok = [c for c in cands if c["status"] == 0 and math.isfinite(c["ms"])]
    best = min(c["ms"] for c in ok)
    t_pick = [c["ms"] for c in ok if c["format"] == fmt and c["tuning"] == tuning]
    assert t_pick, (fmt, tuning)
    RETIMED[k] = {"index": r["index"], "pick": fmt, "regret": t_pick[0] / best - 1.0,
                  "recorded_exhaustive_ms": r["exhaustive_ms"], "fresh_exhaustive_ms": best}
    x.free()
    y.free()
    A.destroy()


def test_fresh_regret_of_the_retimed_matrices(gpu):
    # runs after the parametrized cases above (definition order)
    assert len(RETIMED) == len(_retime_indices())
    rg = np.array([v["regret"] for v in RETIMED.values()])
    print(json.dumps({"retimed": len(rg), "mean_regret": float(rg.mean()), "p90_regret": float(np.quantile(rg, 0.9)),
                      "within_5pct": float(np.mean(rg <= 0.05))}))
    assert rg.mean() <= 0.05
