"""Matrices shared by the GPU parity tests (small enough for the oracle)."""
import functools
import os
import sys

import numpy as np

import matrix_format_performance_amd as mfp

# MFP_CASE_LOG=<path>: also append each case name to that file
LOG = os.environ.get("MFP_CASE_LOG")


def announce(msg: str) -> None:
    """Name the case on stderr (and in MFP_CASE_LOG) before it touches the GPU."""
    print(f"[case] {msg}", file=sys.__stderr__, flush=True)
    if LOG:
        try:
            with open(LOG, "a") as f:
                f.write(msg + "\n")
        except OSError:
            pass


def _edge_rows():
    # empty rows at the start, middle and end; one long row; single entries
    lens = [0, 0, 3, 0, 1, 5000, 0, 2, 64, 65, 0, 1] + [7] * 300 + [0] * 200
    rng = np.random.default_rng(42)
    cols = 6000
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
    ci = np.concatenate([np.sort(rng.choice(cols, n, replace=False)) for n in lens if n] or [[]]).astype(np.uint32)
    va = rng.uniform(-1, 1, ci.size)
    return rp, ci, va, cols


@functools.lru_cache(maxsize=None)
def case(name: str, dtype_str: str):
    """-> (row_ptr, columns, values, cols) as numpy arrays."""
    dtype = np.dtype(dtype_str)
    if name == "stencil2d_64":
        rp, ci, va = mfp.generate(mfp.GenSpec("stencil2d", values="uniform", n=64, seed=1), dtype)
        cols = 64 * 64
    elif name == "stencil3d_20":
        rp, ci, va = mfp.generate(mfp.GenSpec("stencil3d", values="normal", n=20, seed=2), dtype)
        cols = 20 ** 3
    elif name == "powerlaw_20k":
        rp, ci, va = mfp.generate(mfp.GenSpec("powerlaw", values="uniform", rows=20000, nnz_lo=1,
                                              nnz_hi=12000, seed=3), dtype)
        cols = 20000
    elif name == "powerlaw_narrow":
        # power-law row lengths capped at 300: an ELL of it stays small
        rp, ci, va = mfp.generate(mfp.GenSpec("powerlaw", values="uniform", rows=20000, nnz_lo=1,
                                              nnz_hi=300, seed=3), dtype)
        cols = 20000
    elif name == "uniform_band":
        rp, ci, va = mfp.generate(mfp.GenSpec("uniform", values="uniform", rows=30000, nnz_lo=4,
                                              nnz_hi=64, bandwidth=0.02, seed=4), dtype)
        cols = 30000
    elif name == "edge_rows":
        rp, ci, va, cols = _edge_rows()
        va = va.astype(dtype)
    elif name == "one_by_one":
        rp, ci, va, cols = np.array([0, 1], np.uint32), np.array([0], np.uint32), np.array([0.5], dtype), 1
    else:
        raise KeyError(name)
    return rp, ci, np.asarray(va, dtype), cols


CASES = ["stencil2d_64", "stencil3d_20", "powerlaw_20k", "uniform_band", "edge_rows", "one_by_one"]
