"""The C++ benchmark driver (mfp_bench over include/mfp_compat.hpp): the
reference's seven JSON files, labels, byte/flop models and matrix metadata,
for generated matrices and for MatrixMarket files."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle
from gpu_cases import announce

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(HERE, "matrix_format_performance_amd", "mfp_bench")

GPU_LABELS = {"GPU CSR", "GPU CSR (vector)", "GPU CSR-Adaptive", "GPU ELL", "GPU COO", "GPU SCOO",
              "GPU Hybrid (TODO)", "GPU Hybrid (atomic)"}
FILES = ["matrices_info.json"] + [f"{p}{s}.json" for p in ("float", "double")
                                  for s in ("", "_effective_bandwidth", "_computational_throughput")]


def _rocsparse_available():
    """Whether the benchmark-only rocSPARSE slot can open librocsparse.so here."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(HERE, "matrix_format_performance_amd", "librocsparse_slot.so"))
    return bool(lib.mfp_rs_available())


def _run(args, out, timeout=300):
    announce("mfp_bench " + " ".join(args))
    r = subprocess.run([APP, "--out", str(out), *args], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r


def _load(out):
    return {f: json.load(open(os.path.join(out, f))) for f in FILES}


def test_synthetic_run_writes_the_seven_files(gpu, tmp_path):
    specs = ["stencil2d:n=64,values=laplace", "powerlaw:rows=20000,lo=1,hi=12000,seed=3"]
    # one call per record, so bandwidth x elapsed gives back the byte model exactly
    r = _run(["--reps", "1", "--synthetic", ";".join(specs)], tmp_path)
    # compare_results prints "ERROR:" only for a relative squared error above 1e-9
    assert "ERROR" not in r.stderr, r.stderr[-2000:]
    js = _load(tmp_path)
    gen = {specs[0]: gpu.generate(gpu.GenSpec("stencil2d", values="laplace", n=64)),
           specs[1]: gpu.generate(gpu.GenSpec("powerlaw", rows=20000, nnz_lo=1, nnz_hi=12000, seed=3))}
    info = js["matrices_info.json"]
    assert sorted(info) == sorted(specs)
    for name, (rp, ci, _) in gen.items():
        rows, nnz = rp.size - 1, ci.size
        assert info[name] == {"cols": rows, "nnz": int(nnz), "rows": rows}
        width = int(np.diff(rp.astype(np.int64)).max())
        for prec, dt in (("float", np.float32), ("double", np.float64)):
            t = js[f"{prec}.json"][name]
            bw = js[f"{prec}_effective_bandwidth.json"][name]
            fl = js[f"{prec}_computational_throughput.json"][name]
            # the rocSPARSE row (the cuSPARSE slot of main.cpp:194-197) is there
            # exactly when librocsparse.so can be opened
            rs = {"GPU CSR (rocSPARSE)"} if _rocsparse_available() else set()
            assert set(t) == GPU_LABELS | {"CPU CSR", "CPU CSR Parallel"} | rs
            assert set(bw) == set(t) == set(fl)
            assert all(v > 0 for v in t.values())
            csr_b = oracle.reference_bytes(0, dt, rows, nnz)
            models = {  # label -> (bytes, operations) the reference records
                "CPU CSR": (csr_b, 2.0 * nnz), "CPU CSR Parallel": (csr_b, 2.0 * nnz),
                "GPU CSR": (csr_b, 2.0 * nnz), "GPU CSR (vector)": (csr_b, 2.0 * nnz),
                "GPU COO": (oracle.reference_bytes(2, dt, rows, nnz), 2.0 * nnz),
                "GPU ELL": (oracle.reference_bytes(1, dt, rows, nnz, width), 2.0 * ((rows * width) % 2 ** 32)),
                "GPU CSR-Adaptive": (0.0, 0.0), "GPU SCOO": (0.0, 0.0), "GPU Hybrid (atomic)": (0.0, 0.0),
                "GPU Hybrid (TODO)": (1.0, 1.0),
            }
            if rs:  # the reference's cuSPARSE byte model (gpu_matrix_multiplier.cu:245-249)
                models["GPU CSR (rocSPARSE)"] = (csr_b, 2.0 * nnz)
            for label, (b, ops) in models.items():
                assert bw[label] == pytest.approx(b / (t[label] * 1e9), rel=1e-12), (prec, label)
                assert fl[label] == pytest.approx(ops / (t[label] * 1e9), rel=1e-12), (prec, label)


def test_matrix_market_list_and_meta_mode(gpu, tmp_path):
    rng = np.random.default_rng(0)
    # a general real matrix with an empty row, and a pattern matrix
    rows, cols = 50, 40
    ent = sorted({(int(rng.integers(0, rows)), int(rng.integers(0, cols))) for _ in range(300)})
    ent = [(r, c) for r, c in ent if r != 7]
    p1 = tmp_path / "real_general.mtx"
    with open(p1, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% a comment\n")
        f.write(f"{rows} {cols} {len(ent)}\n")
        for r, c in reversed(ent):  # out of order on purpose: the converter sorts
            f.write(f"{r + 1} {c + 1} {rng.uniform(-1, 1):.17g}\n")
    p2 = tmp_path / "pattern.mtx"
    with open(p2, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n3 3 4\n1 1\n2 2\n3 3\n3 1\n")
    lst = tmp_path / "list.txt"
    lst.write_text(f"{p1}\n{p2}\n")
    out = tmp_path / "out"
    out.mkdir()
    r = _run(["--reps", "2", str(lst)], out)
    assert "ERROR" not in r.stderr, r.stderr[-2000:]
    js = _load(out)
    info = js["matrices_info.json"]
    assert info == {"real_general.mtx": {"cols": cols, "nnz": len(ent), "rows": rows},
                    "pattern.mtx": {"cols": 3, "nnz": 4, "rows": 3}}
    assert set(js["double.json"]["pattern.mtx"]) >= GPU_LABELS
    # meta mode: std_deviation of the row lengths, as get_rows_statistics computes it
    out2 = tmp_path / "meta"
    out2.mkdir()
    _run(["--meta-only", str(lst)], out2)
    info2 = json.load(open(out2 / "matrices_info.json"))
    lens = np.bincount([r for r, _ in ent], minlength=rows)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    assert info2["real_general.mtx"]["std_deviation"] == pytest.approx(oracle.rows_statistics(rp)[3], rel=1e-12)
    assert info2["pattern.mtx"]["nnz"] == 4
    assert json.load(open(out2 / "float.json")) is None  # no measurements in this mode


def test_hybrid_sweep_rows(gpu, tmp_path):
    # --hybrid-sweep runs main.cpp:247-279 (switched off in the reference):
    # three ELL widths x three hybrids, each timed once; the host-COO row has
    # its own label and the reference's (1, 1) byte/flop record
    r = _run(["--reps", "1", "--no-cpu", "--hybrid-sweep", "--synthetic", "powerlaw:rows=20000,lo=1,hi=3000,seed=4"],
             tmp_path)
    assert "ERROR" not in r.stderr, r.stderr[-2000:]
    assert r.stdout.count("hybrid sweep:") == 6  # float and double
    assert r.stdout.count("GPU Hybrid (CPU COO TODO)") == 6
    js = _load(tmp_path)
    for prec in ("float", "double"):
        t = js[f"{prec}.json"]["powerlaw:rows=20000,lo=1,hi=3000,seed=4"]
        bw = js[f"{prec}_effective_bandwidth.json"]["powerlaw:rows=20000,lo=1,hi=3000,seed=4"]
        lab = "GPU Hybrid (CPU COO TODO)"
        assert t[lab] > 0 and bw[lab] == pytest.approx(1.0 / (t[lab] * 1e9), rel=1e-12)


def test_symmetric_matrix_is_refused(gpu, tmp_path):
    p = tmp_path / "sym.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n2 2 2\n1 1 1.0\n2 1 2.0\n")
    lst = tmp_path / "list.txt"
    lst.write_text(f"{p}\n")
    announce("mfp_bench symmetric")
    r = subprocess.run([APP, "--out", str(tmp_path), str(lst)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "Only general matrices are supported" in r.stderr
