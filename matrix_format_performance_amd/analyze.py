"""Results analysis over the benchmark driver's JSON files.

The driver (``mfp_bench``, like the reference's ``main.cpp:482-501``) writes
``{float,double}.json`` ({matrix: {label: seconds}}), the matching
``*_effective_bandwidth.json`` / ``*_computational_throughput.json`` files and
``matrices_info.json`` ({matrix: {nnz, rows, cols, std_deviation}}).  This
module turns them into:

- the speedup tables of the write-up (``doc/post.tex:516-530`` and the seven
  tables after it): speedup = t[base] / t[format] per matrix, with base
  "CPU CSR", then the mean and the max over the matrices with nnz above a
  limit (10^4 and 10^5).  This is the statistic of
  ``scripts/analyze.py:25-32`` (per-matrix ratio of the two time columns) and
  ``:68-76`` (rows with nnz > limit, then ``describe()``: mean, max);
- the format-selection table: for every format, the geometric mean over the
  matrices (nnz above a limit) of t[format] / min over the formats of t, i.e.
  the slowdown of always picking that format against the per-matrix best;
- the per-matrix best counts.

No reference code is used; it is read only through its JSON output.
Run ``python -m matrix_format_performance_amd.analyze RESULTS_DIR``.
"""
from __future__ import annotations

import argparse
import gzip
import json
import math
import os
from typing import Dict, Iterable, List, Optional, Sequence

BASE = "CPU CSR"
LIMITS = (10000, 100000)
# the reference's GPU labels (main.cpp:179-245), plus the rocSPARSE slot of
# this project's driver, in the order of the write-up's tables
GPU_FORMATS = ("GPU CSR", "GPU CSR (cuSparse)", "GPU CSR (rocSPARSE)", "GPU CSR (vector)", "GPU CSR-Adaptive",
               "GPU ELL", "GPU COO", "GPU SCOO", "GPU Hybrid (atomic)")

Times = Dict[str, Dict[str, float]]
Info = Dict[str, Dict[str, float]]


def _read_json(path: str):
    # a file may also be stored gzip-compressed as <name>.gz
    if not os.path.exists(path) and os.path.exists(path + ".gz"):
        with gzip.open(path + ".gz", "rt") as f:
            return json.load(f)
    with open(path) as f:
        return json.load(f)


def load(results_dir: str, precision: str = "float", kind: str = "") -> Times:
    """{matrix: {label: value}} from <precision>[_<kind>].json (kind "" = seconds)."""
    return _read_json(os.path.join(results_dir, precision + (f"_{kind}" if kind else "") + ".json"))


def load_info(results_dir: str) -> Info:
    return _read_json(os.path.join(results_dir, "matrices_info.json"))


def _matrices(times: Times, info: Info, limit: float) -> List[str]:
    # the reference merges the time table with matrices_info on the matrix
    # name (an inner join) before filtering on nnz (scripts/analyze.py:20-21)
    return [m for m in times if m in info and info[m]["nnz"] > limit]


def speedups(times: Times, info: Info, label: str, limit: float, base: str = BASE) -> List[float]:
    """t[base] / t[label] for every matrix with nnz > limit that has both."""
    out = []
    for m in _matrices(times, info, limit):
        row = times[m]
        if base in row and label in row:
            t = row[label]
            out.append(row[base] / t if t else math.inf)
    return out


def speedup_table(times: Times, info: Info, labels: Optional[Iterable[str]] = None, limits: Sequence[float] = LIMITS,
                  base: str = BASE) -> Dict[str, Dict[float, dict]]:
    """{label: {limit: {"avg", "max", "count"}}} (the write-up's tables)."""
    if labels is None:
        present = set()
        for row in times.values():
            present |= set(row)
        labels = [l for l in present if l != base]
    table = {}
    for label in labels:
        per = {}
        for lim in limits:
            s = speedups(times, info, label, lim, base)
            per[lim] = {"avg": sum(s) / len(s) if s else math.nan, "max": max(s) if s else math.nan,
                        "count": len(s)}
        table[label] = per
    return table


def regret_table(times: Times, info: Info, formats: Sequence[str], limit: float = 100000) -> Dict[str, dict]:
    """For every format f: the geometric mean over matrices (nnz > limit, every
    format timed) of t[f] / min_g t[g], the median of that ratio and the
    fraction of matrices where f is within 5% of the best."""
    ms = [m for m in _matrices(times, info, limit) if all(f in times[m] for f in formats)]
    out = {}
    for f in formats:
        r = sorted(times[m][f] / min(times[m][g] for g in formats) for m in ms)
        if not r:
            out[f] = {"geomean": math.nan, "median": math.nan, "within_5pct": math.nan, "count": 0}
            continue
        gm = math.exp(sum(math.log(v) for v in r) / len(r))
        med = r[len(r) // 2] if len(r) % 2 else 0.5 * (r[len(r) // 2 - 1] + r[len(r) // 2])
        out[f] = {"geomean": gm, "median": med, "within_5pct": sum(v <= 1.05 for v in r) / len(r),
                  "count": len(r)}
    return out


def best_counts(times: Times, formats: Sequence[str]) -> Dict[str, int]:
    """How often each format is the fastest (matrices that have every format)."""
    cnt = {f: 0 for f in formats}
    for row in times.values():
        if all(f in row for f in formats):
            cnt[min(formats, key=lambda f: row[f])] += 1
    return cnt


def present_formats(times: Times, candidates: Sequence[str] = GPU_FORMATS) -> List[str]:
    """The candidate labels that every matrix of the run has."""
    return [f for f in candidates if times and all(f in row for row in times.values())]


def report(results_dir: str, limits: Sequence[float] = LIMITS, regret_limit: float = 100000) -> dict:
    info = load_info(results_dir)
    out = {}
    for prec in ("float", "double"):
        path = os.path.join(results_dir, f"{prec}.json")
        if not os.path.exists(path) and not os.path.exists(path + ".gz"):
            continue
        times = load(results_dir, prec)
        fmts = present_formats(times)
        out[prec] = {"speedup": speedup_table(times, info, fmts, limits),
                     "regret": regret_table(times, info, fmts, regret_limit),
                     "best_counts": best_counts(times, fmts)}
    return out


def format_report(rep: dict, limits: Sequence[float] = LIMITS) -> str:
    lines = []
    precs = [p for p in ("float", "double") if p in rep]
    labels = list(rep[precs[0]]["speedup"]) if precs else []
    lines.append("Speedup over CPU CSR (avg / max over matrices with nnz > limit)")
    head = f"{'format':24s}" + "".join(f" | {p} nnz>{int(l)}".ljust(22) for p in precs for l in limits)
    lines.append(head)
    for label in labels:
        cells = []
        for p in precs:
            for l in limits:
                c = rep[p]["speedup"].get(label, {}).get(l)
                cells.append(f" | {c['avg']:6.2f} / {c['max']:6.2f} ({c['count']})".ljust(22) if c else " | -".ljust(22))
        lines.append(f"{label:24s}" + "".join(cells))
    for p in precs:
        lines.append("")
        lines.append(f"Always pick one format ({p}): geomean slowdown vs the per-matrix best, nnz > 1e5")
        for f, r in sorted(rep[p]["regret"].items(), key=lambda kv: kv[1]["geomean"]):
            lines.append(f"  {f:24s} {r['geomean']:5.2f}x  median {r['median']:5.2f}x  "
                         f"within 5%: {100 * r['within_5pct']:5.1f}%  ({r['count']} matrices)")
        lines.append(f"Per-matrix best counts ({p}): " +
                     ", ".join(f"{f} {n}" for f, n in sorted(rep[p]["best_counts"].items(), key=lambda kv: -kv[1])))
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("results_dir")
    ap.add_argument("--json", action="store_true", help="print the report as JSON")
    a = ap.parse_args(argv)
    rep = report(a.results_dir)
    if a.json:
        print(json.dumps(rep, indent=1, sort_keys=True, default=str))
    else:
        print(format_report(rep))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
