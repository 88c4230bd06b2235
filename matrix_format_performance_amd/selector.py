"""Learned format selector.

A matrix is described by 64 structure features (``features``, computed on the
device by ``mfp_matrix_features``); an MLP 64 -> 128 -> 128 -> 8 maps them to
a distribution over the eight formats.  The network is trained on the GPU by
the library's own HIP kernels (forward, backward, Adam) against targets made
from timings this package measures (``sweep``: every autotuner candidate of
every format, per matrix), and it infers in FP32 or in BF16 on the matrix
cores.  The pick is the most probable format; its tuning comes from rules
fitted on the same timings (per bin of mean row length and row count, the
tuning with the lowest mean slowdown within the format; a label that lets
the plan size its own work items ranks ahead of a fixed piece count unless
that one wins by more than ``SELF_SIZED_MARGIN``).  The exhaustive
autotuner stays the fallback and the yardstick for the pick's regret.

Command line (``python -m matrix_format_performance_amd.selector``):

* ``sweep --out S.jsonl --count N [--start I] [--max-nnz M] [--min-rows R]
  [--seed K]``: generates matrices I..N-1, records features and candidate
  timings, one JSON line per matrix; rerunning with the same file resumes
  after the last complete line.
* ``train --data S.jsonl [T.jsonl ...] --out W.npz [--hold-every 5]``: holds
  out every fifth record, trains on the others, fits the tuning rules, saves
  both (``Selector.load``) and reports regret on the held-out records.
  The data files may be gzipped (``.jsonl.gz``).
* ``refit-rules --data S.jsonl [...] --out W.npz [--weights V.npz]``: keeps
  the network of V (the shipped one by default), fits the tuning rules again
  on the same training split, saves both and reports as ``train`` does.
"""
from __future__ import annotations

import argparse
import ctypes as C
import gzip
import json
import math
import os
import sys
import time

import numpy as np

from . import _lib as L
from . import engine as E

NI, NH, NO = 64, 128, 8
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_mi355x.npz")
FORMAT_ORDER = [E.FORMAT_NAMES[i] for i in range(NO)]


def features(matrix: "E.DeviceMatrix") -> np.ndarray:
    """The 64 structure features of a device matrix (float32)."""
    out = np.empty(NI, np.float32)
    L.call("mfp_matrix_features", matrix.handle, out.ctypes.data)
    return out


class MLP:
    """The selector network; parameters live on the device."""

    def __init__(self, seed: int = 0):
        E.lib()
        h = C.c_void_p()
        L.call("mfp_mlp_create", C.byref(h), seed)
        self.handle = h.value
        self.nparam = L.load().mfp_mlp_param_count()

    @property
    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, np.float32)
        L.call("mfp_mlp_get_params", self.handle, out.ctypes.data)
        return out

    @params.setter
    def params(self, theta) -> None:
        theta = np.ascontiguousarray(theta, np.float32)
        if theta.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {theta.size}")
        L.call("mfp_mlp_set_params", self.handle, theta.ctypes.data)

    @staticmethod
    def _xp(X, P=None):
        X = np.ascontiguousarray(X, np.float32)
        if X.ndim != 2 or X.shape[1] != NI or X.shape[0] == 0:
            raise ValueError(f"X must be n x {NI} with n > 0")
        if P is not None:
            P = np.ascontiguousarray(P, np.float32)
            if P.shape != (X.shape[0], NO):
                raise ValueError(f"P must be {X.shape[0]} x {NO}")
        return X, P

    def loss_grad(self, X, P):
        X, P = self._xp(X, P)
        loss = C.c_double()
        g = np.empty(self.nparam, np.float32)
        L.call("mfp_mlp_loss_grad", self.handle, X.ctypes.data, P.ctypes.data, X.shape[0], C.byref(loss),
               g.ctypes.data)
        return loss.value, g

    def train_step(self, X, P, lr: float) -> float:
        X, P = self._xp(X, P)
        loss = C.c_double()
        L.call("mfp_mlp_train_step", self.handle, X.ctypes.data, P.ctypes.data, X.shape[0], lr, C.byref(loss))
        return loss.value

    def predict(self, X, bf16: bool = False) -> np.ndarray:
        X, _ = self._xp(X)
        out = np.empty((X.shape[0], NO), np.float32)
        L.call("mfp_mlp_predict", self.handle, X.ctypes.data, X.shape[0], int(bool(bf16)), out.ctypes.data)
        return out

    def destroy(self):
        if getattr(self, "handle", None):
            L.load().mfp_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---- targets and regret -------------------------------------------------------

def format_times(record: dict, which: str = "best") -> np.ndarray:
    """Per-format time (ms) of a sweep record: ``best`` over the format's
    candidates or its ``default`` candidate; +inf where none could be built."""
    key = "best_ms" if which == "best" else "default_ms"
    return np.array([record[key].get(f, math.inf) for f in FORMAT_ORDER], np.float64)


def soft_targets(times: np.ndarray, tau: float = 0.05) -> np.ndarray:
    """Target distribution over formats: p_f proportional to
    exp(-(t_f / t_best - 1) / tau), zero for formats that could not run."""
    times = np.atleast_2d(times)
    best = np.min(times, axis=1, keepdims=True)
    z = -(times / best - 1.0) / tau
    z[~np.isfinite(times)] = -np.inf
    z -= z.max(axis=1, keepdims=True)
    p = np.exp(z)
    return (p / p.sum(axis=1, keepdims=True)).astype(np.float32)


def regret(times_pick: np.ndarray, times_best: np.ndarray) -> np.ndarray:
    """t(pick) / t(exhaustive best) - 1 per matrix."""
    return np.asarray(times_pick, np.float64) / np.asarray(times_best, np.float64) - 1.0


def evaluate(records, probs: np.ndarray) -> dict:
    """Regrets of the network's pick against the exhaustive best (every
    candidate of every format) and of the best single fixed format."""
    best_t = np.array([r["exhaustive_ms"] for r in records])
    tb = np.stack([format_times(r, "best") for r in records])
    td = np.stack([format_times(r, "default") for r in records])
    pick = np.argmax(probs, axis=1)
    rows = np.arange(len(records))
    r_default = regret(td[rows, pick], best_t)
    r_tuned = regret(tb[rows, pick], best_t)
    fixed = {}
    for j, f in enumerate(FORMAT_ORDER):
        rr = regret(td[:, j], best_t)
        fixed[f] = float(np.mean(rr)) if np.all(np.isfinite(rr)) else math.inf
    best_fixed = min(fixed, key=fixed.get)
    return {
        "count": len(records),
        "mean_regret": float(np.mean(r_default)),
        "geomean_slowdown": float(np.exp(np.mean(np.log1p(r_default)))),
        "median_regret": float(np.median(r_default)),
        "p90_regret": float(np.quantile(r_default, 0.9)),
        "within_5pct": float(np.mean(r_default <= 0.05)),
        "mean_regret_tuned_format": float(np.mean(r_tuned)),
        "format_accuracy": float(np.mean(pick == np.argmin(tb, axis=1))),
        "fixed_format_mean_regret": fixed,
        "best_fixed_format": best_fixed,
        "best_fixed_mean_regret": fixed[best_fixed],
        "picks": {f: int(np.sum(pick == j)) for j, f in enumerate(FORMAT_ORDER)},
    }


def train(X, P, epochs: int = 300, batch: int = 128, lr: float = 2e-3, seed: int = 0, mlp: MLP = None,
          log=None) -> MLP:
    """Adam on minibatches of (features, target distribution), on the GPU."""
    X = np.ascontiguousarray(X, np.float32)
    P = np.ascontiguousarray(P, np.float32)
    net = mlp or MLP(seed)
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    for ep in range(epochs):
        order = rng.permutation(n)
        tot = 0.0
        for s in range(0, n, batch):
            idx = order[s:s + batch]
            tot += net.train_step(X[idx], P[idx], lr) * idx.size
        if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
            log(f"epoch {ep} loss {tot / n:.4f}")
    return net


# ---- the sweep ------------------------------------------------------------------

def sample_spec(index: int, seed: int = 0, max_nnz: int = 1 << 24, min_rows: int = 10_000,
                max_rows: int = 4 << 20) -> "E.GenSpec":
    """Matrix ``index`` of the selector dataset: families 2D/3D stencils,
    uniform nnz/row and power law; rows log-uniform in [min_rows, max_rows],
    mean nnz/row about 2-200, bandwidth fraction 0.001-1; rows reduced so that
    nnz stays near or below max_nnz."""
    rng = np.random.default_rng([seed, index])
    fam = ["stencil2d", "stencil3d", "uniform", "powerlaw", "uniform", "powerlaw"][index % 6]
    rows = int(math.exp(rng.uniform(math.log(min_rows), math.log(max_rows))))
    vals = ["uniform", "normal"][index % 2]
    if fam == "stencil2d":
        rows = min(rows, max_nnz // 5)
        return E.GenSpec("stencil2d", values=vals, n=max(8, int(math.sqrt(rows))), seed=index)
    if fam == "stencil3d":
        rows = min(rows, max_nnz // 27)
        return E.GenSpec("stencil3d", values=vals, n=max(4, int(round(rows ** (1 / 3)))), seed=index)
    bw = float(math.exp(rng.uniform(math.log(1e-3), 0.0)))
    if fam == "uniform":
        mean = math.exp(rng.uniform(math.log(2), math.log(200)))
        spread = rng.uniform(0.0, 1.0)
        rows = max(min_rows // 4, min(rows, int(max_nnz / mean)))
        lo = max(1, int(mean * (1 - spread)))
        hi = max(lo, int(mean * (1 + spread)))
        return E.GenSpec("uniform", values=vals, rows=rows, nnz_lo=lo, nnz_hi=hi, bandwidth=bw, seed=index)
    alpha = float(rng.uniform(1.5, 3.0))
    lo = int(math.exp(rng.uniform(0.0, math.log(64))))
    hi = int(min(rows, lo * 10 ** rng.uniform(1.0, 4.0)))
    # E[min(X, hi)] for X with density ~ x^-alpha on [lo, inf): lo + int_lo^hi (lo/x)^(alpha-1) dx
    if abs(alpha - 2.0) < 1e-9:
        mean = lo * (1 + math.log(hi / lo))
    else:
        mean = lo + lo / (2.0 - alpha) * ((hi / lo) ** (2.0 - alpha) - 1.0)
    rows = max(min_rows // 4, min(rows, int(max_nnz / max(mean, 1.0))))
    hi = max(lo, min(hi, rows))
    return E.GenSpec("powerlaw", values=vals, rows=rows, nnz_lo=lo, nnz_hi=hi, alpha=alpha, bandwidth=bw,
                     seed=index)


def row_length_marks(row_ptr) -> dict:
    """The row-length marks the autotuner's hybrid widths come from: the
    rounded-up mean and the 0.75 / 0.9 / 0.99 quantiles (element
    floor(q (rows - 1)) of the sorted lengths)."""
    lens = np.diff(np.asarray(row_ptr, np.int64))
    n = lens.size
    if n == 0:
        return {}
    srt = np.sort(lens)
    marks = {"avg": int(-(-int(srt.sum()) // n))}
    for name, q in (("q75", 0.75), ("q90", 0.9), ("q99", 0.99)):
        marks[name] = int(srt[min(n - 1, int(q * (n - 1)))])
    return marks


def tuning_label(fmt: str, t: dict, marks: dict, dtype=np.float32) -> str:
    """A name for a candidate's tuning that means the same on every matrix."""
    if fmt == "csr_vector":
        return f"lanes{t['lanes']}"
    if fmt == "csr_adaptive":
        base = E.DEFAULT_NNZ_PER_WG[np.dtype(dtype)]
        return {base // 2: "W/2", base: "W", base * 2: "2W"}.get(t["nnz_per_wg"], f"W{t['nnz_per_wg']}")
    if fmt == "scoo":
        if t["flags"] & L.MFP_TUNE_SCOO_COLSORT:
            th = "t1024" if t.get("scoo_threads") == 1024 else ""
            return f"cs{t['scoo_slice_size']}p{t.get('scoo_pieces', 0)}{th}"
        return "default" if t["flags"] & L.MFP_TUNE_IS_DEFAULT else f"ss{t['scoo_slice_size']}"
    if fmt == "ell":
        return "packed" if t["flags"] & L.MFP_TUNE_ELL_PACKED else "default"
    if fmt in ("hyb", "hyb_atomic"):
        w = t["hyb_width"]
        if w == 0:
            return "rule"
        for name in ("avg", "q75", "q90", "q99"):
            if marks.get(name) == w:
                return name
        return f"w{w}"
    return "default"


def measure(matrix: "E.DeviceMatrix", reps: int = 3, flush: bool = True, row_ptr=None) -> dict:
    """Features and exhaustive candidate timings of one device matrix."""
    vt = matrix.vector_dtype
    x = E.DeviceArray.from_host(E.generate_vector(1, matrix.cols, vt))
    y = E.DeviceArray(matrix.rows, vt)
    t0 = time.perf_counter()
    feats = features(matrix)
    t_feat = time.perf_counter() - t0
    t0 = time.perf_counter()
    cands = E.autotune(matrix, x, y, reps=reps, flush=flush)
    t_tune = time.perf_counter() - t0
    if row_ptr is None:
        row_ptr = matrix.download()[0]
    marks = row_length_marks(row_ptr)
    best_ms, default_ms, listed = {}, {}, []
    for c in cands:
        f = c["format"]
        ok = c["status"] == 0 and math.isfinite(c["ms"])
        listed.append([f, tuning_label(f, c["tuning"], marks, vt), c["ms"] if ok else None,
                       bool(c["tuning"]["flags"] & L.MFP_TUNE_IS_DEFAULT)])
        if not ok:
            continue
        best_ms[f] = min(best_ms.get(f, math.inf), c["ms"])
        if c["tuning"]["flags"] & L.MFP_TUNE_IS_DEFAULT:
            default_ms[f] = c["ms"]
    x.free()
    y.free()
    return {
        "features": [float(v) for v in feats],
        "best_ms": best_ms,
        "default_ms": default_ms,
        "exhaustive_ms": min(best_ms.values()),
        "candidates": len(cands),
        "cands": listed,
        "marks": marks,
        "feature_s": t_feat,
        "autotune_s": t_tune,
    }


# ---- tuning rules: the tuning of the picked format --------------------------------

def tuning_keys(feats) -> list:
    """Lookup keys from the most to the least specific: (floor log2(1 + mean
    row length), floor log2(rows + 1), column spread bin), then the first two
    alone, then the mean bin, then none.  The spread bin comes from feature 58
    (mean row span over cols): 0 below 0.05 (banded), 1 below 0.3, else 2;
    it decides how far column sorting can bring a slice's x reads together."""
    m = int(round(feats[3] * 16 * 1e6)) // 1000000
    r = int(round(feats[0] * 32 * 1e6)) // 1000000
    s = 0 if feats[58] < 0.05 else (1 if feats[58] < 0.3 else 2)
    return [f"{m},{r},{s}", f"{m},{r}", f"{m}", ""]


# A label whose plan sizes its own work items from the matrix and the CU
# count (the column-sorted SCOO with its pieces chosen by the plan, "p0")
# ranks ahead of a fixed piece count unless that one is better by more than
# this mean slowdown.  A fixed count fits the training matrices of a key but
# not every matrix in it: on the 1M-row power-law matrix 2 pieces per
# 16,384-row slice leave 128 work items for 256 CUs (53.2 against 42.4 us),
# while 6 training matrices of its key ranked 2 pieces first by 0.2%.
# Chosen on a validation split of the training records (the margins 0 to
# 0.15 tried; 0.1 and above gave the lowest validation regret).
SELF_SIZED_MARGIN = 0.1


def _self_sized(label: str) -> bool:
    return label.startswith("cs") and "p0" in label


def fit_tuning_rules(records, min_count: int = 3, self_sized_margin: float = SELF_SIZED_MARGIN) -> dict:
    """For every format and key, tuning labels ordered by their mean slowdown
    against the format's best tuning on the training matrices (a label that
    failed to build counts as 10x slower), self-sized labels credited with
    self_sized_margin."""
    rules = {}
    for f in FORMAT_ORDER:
        agg = {}
        for r in records:
            cs = [c for c in r.get("cands", []) if c[0] == f]
            ok = [c[2] for c in cs if c[2] is not None]
            if not ok:
                continue
            bestf = min(ok)
            for key in tuning_keys(r["features"]):
                for c in cs:
                    agg.setdefault(key, {}).setdefault(c[1], []).append(
                        c[2] / bestf - 1.0 if c[2] is not None else 10.0)
        table = {}
        for key, labels in agg.items():
            n = max(len(v) for v in labels.values())
            if n < min_count and key != "":
                continue
            table[key] = [lab for lab, _ in sorted(
                labels.items(),
                key=lambda kv: (float(np.mean(kv[1])) - (self_sized_margin if _self_sized(kv[0]) else 0.0), kv[0]))]
        rules[f] = table
    return rules


def choose_label(rules: dict, fmt: str, feats, available) -> str:
    """The best-ranked label of the most specific key that the matrix has."""
    table = rules.get(fmt, {})
    for key in tuning_keys(feats):
        for lab in table.get(key, []):
            if lab in available:
                return lab
    return "default" if "default" in available else sorted(available)[0]


class Selector:
    """The trained network plus the tuning rules: picks (format, tuning) for
    a matrix from its features, without timing anything."""

    def __init__(self, net: MLP, rules: dict):
        self.net, self.rules = net, rules

    def save(self, path: str) -> None:
        np.savez(path, params=self.net.params, rules=np.array(json.dumps(self.rules)))

    @classmethod
    def default(cls) -> "Selector":
        """The selector trained on this package's 2,400-matrix MI355X sweep
        (selector_mi355x.npz next to this file)."""
        return cls.load(DEFAULT_WEIGHTS)

    @classmethod
    def load(cls, path: str) -> "Selector":
        with np.load(path, allow_pickle=False) as z:
            net = MLP(0)
            net.params = z["params"]
            return cls(net, json.loads(str(z["rules"])))

    def rank(self, feats, bf16: bool = True) -> list:
        probs = self.net.predict(np.asarray(feats, np.float32)[None, :], bf16=bf16)[0]
        return [FORMAT_ORDER[j] for j in np.argsort(-probs, kind="stable")]

    def pick(self, matrix: "E.DeviceMatrix", bf16: bool = True, row_ptr=None):
        """-> (format, tuning dict) for a device matrix."""
        feats = features(matrix)
        if row_ptr is None:
            row_ptr = matrix.download()[0]
        marks = row_length_marks(row_ptr)
        for fmt in self.rank(feats, bf16):
            cands = E.autotune_candidates(matrix, [fmt])
            if not cands:
                continue
            labelled = {tuning_label(fmt, c["tuning"], marks, matrix.vector_dtype): c["tuning"] for c in cands}
            return fmt, labelled[choose_label(self.rules, fmt, feats, labelled)]
        raise RuntimeError("no format has candidates for this matrix")


def record_time(record: dict, fmt: str, label: str):
    for c in record.get("cands", []):
        if c[0] == fmt and c[1] == label:
            return c[2] if c[2] is not None else math.inf
    return math.inf


def evaluate_with_rules(records, probs: np.ndarray, rules: dict) -> dict:
    """Regret of the network's format with the rules' tuning."""
    pick = np.argmax(probs, axis=1)
    best_t = np.array([r["exhaustive_ms"] for r in records])
    t = []
    for r, j in zip(records, pick):
        f = FORMAT_ORDER[j]
        avail = {c[1] for c in r["cands"] if c[0] == f}
        t.append(record_time(r, f, choose_label(rules, f, r["features"], avail)) if avail else math.inf)
    rg = regret(np.array(t), best_t)
    return {"mean_regret": float(np.mean(rg)), "median_regret": float(np.median(rg)),
            "geomean_slowdown": float(np.exp(np.mean(np.log1p(rg)))), "p90_regret": float(np.quantile(rg, 0.9)),
            "within_5pct": float(np.mean(rg <= 0.05))}


def _done_indices(path: str) -> set:
    done = set()
    if not os.path.exists(path):
        return done
    with open(path) as f:
        for line in f:
            try:
                done.add(json.loads(line)["index"])
            except (ValueError, KeyError):
                break  # a torn last line: its matrix is measured again
    return done


def sweep(out: str, count: int, seed: int = 0, max_nnz: int = 1 << 24, reps: int = 3, start: int = 0,
          deadline_s: float = None, log=print, min_rows: int = 10_000) -> int:
    """Measure matrices start..count-1 into the JSONL file ``out``, skipping
    those it already holds; returns how many were measured in this call."""
    done = _done_indices(out)
    t_start = time.perf_counter()
    measured = 0
    with open(out, "a") as f:
        for i in range(start, count):
            if i in done:
                continue
            if deadline_s is not None and time.perf_counter() - t_start > deadline_s:
                log(f"deadline reached after {measured} matrices")
                break
            spec = sample_spec(i, seed, max_nnz, min_rows)
            rp, ci, va = E.generate(spec, np.float32)
            cols = {"stencil2d": spec.n * spec.n, "stencil3d": spec.n ** 3}.get(spec.family, rp.size - 1)
            A = E.DeviceMatrix.from_csr(rp, ci, va, cols)
            rec = measure(A, reps=reps, row_ptr=rp)
            A.destroy()
            rec.update({"index": i, "family": spec.family, "rows": int(rp.size - 1), "cols": int(cols),
                        "nnz": int(ci.size), "spec": {k: getattr(spec, k) for k in
                                                      ("n", "rows", "nnz_lo", "nnz_hi", "alpha", "bandwidth")}})
            f.write(json.dumps(rec) + "\n")
            f.flush()
            measured += 1
            if log and measured % 25 == 0:
                log(f"{measured} measured (index {i}, {time.perf_counter() - t_start:.0f} s)")
    return measured


def load_records(path: str) -> list:
    recs = {}
    with (gzip.open(path, "rt") if path.endswith(".gz") else open(path)) as f:
        for line in f:
            try:
                r = json.loads(line)
            except ValueError:
                break
            recs[r["index"]] = r
    return [recs[k] for k in sorted(recs)]


def rule_times(records, rules) -> np.ndarray:
    """Per-format time with the tuning the rules choose (inf where none)."""
    out = np.full((len(records), NO), math.inf)
    for i, r in enumerate(records):
        for j, f in enumerate(FORMAT_ORDER):
            avail = {c[1] for c in r.get("cands", []) if c[0] == f}
            if avail:
                out[i, j] = record_time(r, f, choose_label(rules, f, r["features"], avail))
    return out


def dataset(records, rules=None, tau: float = 0.05):
    """Features and target distributions: from each format's best time, or,
    given rules, from the time of the tuning the rules pick (what a pick of
    that format would get)."""
    X = np.array([r["features"] for r in records], np.float32)
    T = rule_times(records, rules) if rules is not None else np.stack([format_times(r, "best") for r in records])
    return X, soft_targets(T, tau)


def fixed_format_baseline(train_records, test_records, rules=None) -> dict:
    """The method without a selector: the one format with the lowest mean
    regret on the training matrices (default tuning, or the rules' tuning),
    applied to every test matrix."""
    def times(recs):
        return rule_times(recs, rules) if rules is not None else np.stack([format_times(r, "default") for r in recs])
    btr = np.array([r["exhaustive_ms"] for r in train_records])
    ttr = times(train_records)
    means = [float(np.mean(ttr[:, j] / btr - 1)) if np.all(np.isfinite(ttr[:, j])) else math.inf for j in range(NO)]
    j = int(np.argmin(means))
    bte = np.array([r["exhaustive_ms"] for r in test_records])
    rg = times(test_records)[:, j] / bte - 1
    return {"format": FORMAT_ORDER[j], "mean_regret": float(np.mean(rg)),
            "geomean_slowdown": float(np.exp(np.mean(np.log1p(rg)))), "within_5pct": float(np.mean(rg <= 0.05))}


def train_selector(records, tau=0.05, epochs=300, lr=2e-3, batch=128, seed=0, log=None,
                   self_sized_margin=SELF_SIZED_MARGIN):
    """Rules from the records, then the network on rule-based targets."""
    rules = fit_tuning_rules(records, self_sized_margin=self_sized_margin)
    X, P = dataset(records, rules, tau)
    net = train(X, P, epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    return Selector(net, rules)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m matrix_format_performance_amd.selector")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--out", required=True)
    s.add_argument("--count", type=int, default=2000)
    s.add_argument("--start", type=int, default=0)
    s.add_argument("--seed", type=int, default=0)
    s.add_argument("--max-nnz", type=int, default=1 << 24)
    s.add_argument("--min-rows", type=int, default=10_000)
    s.add_argument("--reps", type=int, default=3)
    s.add_argument("--deadline", type=float, default=None, help="stop after this many seconds")
    t = sub.add_parser("train")
    t.add_argument("--data", required=True, nargs="+", help="sweep JSONL files (records joined in order)")
    t.add_argument("--out", required=True)
    t.add_argument("--hold-every", type=int, default=5,
                   help="hold out every n-th record (by position) for the evaluation; 0 = none")
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--report", default=None, help="also write the evaluation JSON here")
    t.add_argument("--tau", type=float, default=0.05, help="target temperature (relative slowdown)")
    t.add_argument("--tune", action="store_true", help="choose tau and epochs on a validation split")
    t.add_argument("--self-sized-margin", type=float, default=SELF_SIZED_MARGIN)
    f = sub.add_parser("refit-rules", help="fit a trained selector's tuning rules again on sweep records, "
                                           "keeping its network, and evaluate it as train does")
    f.add_argument("--weights", default=DEFAULT_WEIGHTS)
    f.add_argument("--data", required=True, nargs="+", help="sweep JSONL files (records joined in order)")
    f.add_argument("--out", required=True)
    f.add_argument("--hold-every", type=int, default=5)
    f.add_argument("--self-sized-margin", type=float, default=SELF_SIZED_MARGIN)
    f.add_argument("--report", default=None, help="also write the evaluation JSON here")
    a = ap.parse_args(argv)
    E.init()
    if a.cmd == "sweep":
        n = sweep(a.out, a.count, a.seed, a.max_nnz, a.reps, a.start, a.deadline,
                  log=lambda m: print(m, flush=True), min_rows=a.min_rows)
        print(json.dumps({"measured": n, "total": len(_done_indices(a.out))}))
        return 0
    recs = [r for path in a.data for r in load_records(path)]
    k = a.hold_every
    tr = [r for i, r in enumerate(recs) if not (k and i % k == k - 1)]
    te = [r for i, r in enumerate(recs) if k and i % k == k - 1]
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.cmd == "refit-rules":
        sl = Selector.load(a.weights)
        sl.rules = fit_tuning_rules(tr, self_sized_margin=a.self_sized_margin)
        sl.save(a.out)
        cfg = {"refit_rules_of": os.path.basename(a.weights), "train": len(tr), "held_out": len(te),
               "self_sized_margin": a.self_sized_margin}
        if a.report and os.path.exists(a.report):
            with open(a.report) as fh:  # the network's own training settings, from its report
                cfg["network"] = json.load(fh).get("config")
        return _report(a, sl, tr, te, cfg, [])
    # hyper-parameters on a validation split of the training matrices
    grid = [(tau, ep) for tau in (0.02, 0.05, 0.1, 0.2) for ep in (150, 300, 600)] if a.tune else [(a.tau, a.epochs)]
    choice, tuning_log = grid[0], []
    if len(grid) > 1:
        nv = len(tr) // 5
        fit, val = tr[:-nv], tr[-nv:]
        best = math.inf
        for tau, ep in grid:
            sl = train_selector(fit, tau, ep, a.lr, seed=a.seed)
            Xv = np.array([r["features"] for r in val], np.float32)
            ev = evaluate_with_rules(val, sl.net.predict(Xv, bf16=True), sl.rules)
            sl.net.destroy()
            tuning_log.append({"tau": tau, "epochs": ep, **ev})
            log(f"tau={tau} epochs={ep}: validation mean regret {ev['mean_regret']:.4f}")
            if ev["mean_regret"] < best:
                best, choice = ev["mean_regret"], (tau, ep)
    tau, ep = choice
    sl = train_selector(tr, tau, ep, a.lr, seed=a.seed, log=log, self_sized_margin=a.self_sized_margin)
    sl.save(a.out)
    return _report(a, sl, tr, te, {"tau": tau, "epochs": ep, "lr": a.lr, "train": len(tr), "held_out": len(te),
                                   "self_sized_margin": a.self_sized_margin}, tuning_log)


def _report(a, sl, tr, te, config, tuning_log) -> int:
    """The evaluation JSON of a selector on its training and held-out records."""
    Xtr = np.array([r["features"] for r in tr], np.float32)
    rep = {"config": config,
           "validation": tuning_log,
           "train_with_rules": evaluate_with_rules(tr, sl.net.predict(Xtr, bf16=True), sl.rules)}
    if te:
        Xt = np.array([r["features"] for r in te], np.float32)
        for name, bf in (("held_out_fp32", False), ("held_out_bf16", True)):
            q = sl.net.predict(Xt, bf16=bf)
            rep[name] = evaluate_with_rules(te, q, sl.rules)
            rep[name]["format_accuracy"] = evaluate(te, q)["format_accuracy"]
        rep["held_out_default_tuning"] = evaluate(te, sl.net.predict(Xt, bf16=True))
        rep["no_selector_fixed_format_default_tuning"] = fixed_format_baseline(tr, te)
        rep["no_selector_fixed_format_rules_tuning"] = fixed_format_baseline(tr, te, sl.rules)
        rep["cost_s"] = {"features_mean": float(np.mean([r["feature_s"] for r in te])),
                         "exhaustive_autotune_mean": float(np.mean([r["autotune_s"] for r in te]))}
        t0 = time.perf_counter()
        for _ in range(20):
            sl.net.predict(Xt[:1], bf16=True)
        rep["cost_s"]["inference_one_matrix"] = (time.perf_counter() - t0) / 20
    print(json.dumps(rep, indent=1))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
