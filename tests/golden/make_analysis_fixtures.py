"""Write the fixtures of tests/test_analyze.py from the reference tree.

- reference_results_scoo/{float,double,matrices_info}.json.gz: the reference's
  committed benchmark output results/scoo/ (RTX 2080 run behind the write-up's
  tables), copied as data, gzip-compressed.
- post_speedup_tables.json: the eight published speedup tables of
  doc/post.tex (avg / max over matrices with nnz > 1e4 and > 1e5, float and
  double), parsed from the LaTeX source.

Run from the repository root with the reference at /root/reference.
"""
import gzip
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

# (label in the JSON files, first line of the table in doc/post.tex)
TABLES = [("GPU CSR", 516), ("GPU CSR (cuSparse)", 536), ("GPU CSR (vector)", 713), ("GPU CSR-Adaptive", 994),
          ("GPU ELL", 1224), ("GPU COO", 1280), ("GPU SCOO", 1318), ("GPU Hybrid (atomic)", 1386)]


def main():
    out_dir = os.path.join(HERE, "reference_results_scoo")
    os.makedirs(out_dir, exist_ok=True)
    for name in ("float.json", "double.json", "matrices_info.json"):
        with open(os.path.join(REF, "results", "scoo", name), "rb") as f, \
                gzip.GzipFile(os.path.join(out_dir, name + ".gz"), "wb", mtime=0) as g:
            shutil.copyfileobj(f, g)
    lines = open(os.path.join(REF, "doc", "post.tex")).read().split("\n")
    tables = {}
    for label, first in TABLES:
        rows = {}
        for ln in range(first - 1, first + 14):
            m = re.match(r"\s*(10000|100000)\s*&\s*([\d.]+)\s*&\s*([\d.]+)\s*&\s*([\d.]+)\s*&\s*([\d.]+)", lines[ln])
            if m:
                rows[m.group(1)] = {"float": {"avg": float(m.group(2)), "max": float(m.group(3))},
                                    "double": {"avg": float(m.group(4)), "max": float(m.group(5))},
                                    "source": f"doc/post.tex:{ln + 1}"}
        assert set(rows) == {"10000", "100000"}, label
        tables[label] = rows
    with open(os.path.join(HERE, "post_speedup_tables.json"), "w") as f:
        json.dump(tables, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
