#!/usr/bin/env python3
"""SpMV benchmark on one MI355X.

One step = one SpMV y = A x over the whole synthetic matrix, inputs resident
in HBM.  Before every step (warm-up and timed) a 1 GiB read evicts the
256 MB Infinity Cache, outside the event pair, so each timed SpMV reads the
matrix from HBM (a read leaves no dirty lines for the SpMV to write back).
The step is timed with a hipEvent pair on the stream the kernels run on.

value = algorithmic bytes / mean step time, in GB/s, where
  B_alg = nnz*(s_v + 4) + (rows + 1)*4 + cols*s_x + rows*s_y
(the CSR compulsory traffic; padding, COO row indices and repeated x reads
count as overhead, not credit).  The plan measured is the fastest autotuner
candidate (every format and its swept tunings, each timed the same way as
the steps), reported in config.format / config.tuning, with the best
candidate of every format in the "formats" table.

roofline: the same flushed steps again, timed per kernel of the SpMV
(mfp_plan_kernel_times: every FP32 / FP64 kernel records its own start and
end); the dominant kernel is the one with the longest mean time.  achieved =
its share of B_alg (kernel_bytes: all of it for a lone kernel; the matrix and
x reads for the main kernel of the column-sorted SCOO, whose finish writes
y; ...) / that mean.  step_achieved / step_frac are B_alg over the whole
step's time (value).  traffic = the dominant kernel's HBM bytes per launch
from rocprofv3 FETCH_SIZE / WRITE_SIZE (profiles/traffic_*.json).

precision: best FP32 form against BF16 / FP16 value storage on the 3D
27-point stencil and the power-law matrix with N(0, 1) values, with errors
against FP64; GB/s there are effective rates (B_alg / time).  f64: the
power-law workload in FP64 (fastest candidate, same step protocol).

learned_pick: the shipped selector's pick on this matrix against the
exhaustive best, and its held-out regret over the committed sweep.

cpu_baseline: the oracle's restatement of the reference CPU path
(cpu_matrix_multiplier.cpp), timed on this host on the same matrix with the
affinity mask's threads (the analog of hardware_concurrency()), capped by
OMP_NUM_THREADS; the CPU model and counts are recorded.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "SpMV GB/s & % HBM roofline per format; learned-pick slowdown vs oracle"

WORKLOADS = {
    # name: (generator workload, description)
    "powerlaw_1m": "power-law nnz/row (Pareto density x^-2, floor, 1..50000), 1,048,576 rows, "
                   "uniform distinct columns, values and x U[-1,1)",
    "stencil3d_128": "3D 27-point stencil 128^3 (2,097,152 rows, 55,742,968 nnz), values N(0,1)",
    "stencil2d_256": "2D 5-point Laplacian 256^2 (65,536 rows, 326,656 nnz)",
    "batch512": "512 matrices, rows log-uniform in [1k, 200k], 2D/3D stencils, uniform 4-64 nnz/row and "
                "power-law alpha in [1.5, 3] (max 10,000 per row), FP32, k = 1, 4, 16 right-hand sides, "
                "assembled into one block-diagonal matrix; each step replays one HIP graph",
}


def run_batch(mfp, args, timer, production=False):
    """One step = one replay of the captured persistent launch over all 512
    matrices (k right-hand sides); value = algorithmic bytes / mean step.
    production: build only the two batches whose forms the full study keeps
    (column-sorted SCOO for k = 1 and 4, CSR-stream SpMM for k = 16; the
    first batch's k = 1 form is CSR-Adaptive), time every form they offer,
    and return the record instead of printing it."""
    copy_gbs = measured_copy_gbs(mfp, timer)
    log(f"[bench] measured 16-byte copy: {copy_gbs:.1f} GB/s")
    t0 = time.time()
    members = mfp.engine.batch_workload()
    dms = [mfp.DeviceMatrix.from_csr(rp, ci, va) for rp, ci, va in members]
    # lanes per row for the k > 1 launch: the mean row length rounded up to a
    # power of two for every row (or half that: wide X rows fill the lanes
    # with fewer entries), each row's own length (binned, lanes 0), or the
    # column-sorted SpMM (lanes 1).  By default all are timed for each k and
    # the fastest is reported.
    tot_rows = sum(m[0].size - 1 for m in members)
    tot_nnz = sum(m[1].size for m in members)
    mean_lanes = int(min(64, max(2, 1 << max(0, int(np.ceil(np.log2(max(1.0, tot_nnz / tot_rows))))))))
    modes = {"mean": mean_lanes, "binned": 0, "colsort": 1, "stream": mfp._lib.MFP_BATCH_STREAM}
    if mean_lanes >= 4:
        modes["half"] = mean_lanes // 2
    if production:
        modes = {"colsort": 1, "stream": mfp._lib.MFP_BATCH_STREAM}
    elif args.batch_lanes == mfp._lib.MFP_BATCH_STREAM:
        modes = {"stream": mfp._lib.MFP_BATCH_STREAM}
    elif args.batch_lanes > 1:
        modes = {f"lanes{args.batch_lanes}": args.batch_lanes}
    elif args.batch_lanes == 0:
        modes = {"binned": 0}
    elif args.batch_lanes == 1:
        modes = {"colsort": 1}
    # k = 1: CSR-Adaptive (the first batch) or the column-sorted SCOO (the last)
    # (8,192-row slices, one item per slice: 2.6% faster than 16,384 rows in
    # two pieces over 33 interleaved flushed launches each)
    cs = {"flags": mfp._lib.MFP_TUNE_SCOO_COLSORT, "scoo_slice_size": 8192, "scoo_pieces": 1,
          "scoo_threads": 1024}
    names = list(modes)
    # the first batch keeps CSR-Adaptive for k = 1, the second (or an extra
    # one) gets the column-sorted SCOO for k = 1
    batches = {n: mfp.Batch(dms, lanes=modes[n], **(cs if i == 1 else {})) for i, n in enumerate(names)}
    if len(names) == 1:
        batches["k1_colsort"] = mfp.Batch(dms, lanes=modes[names[0]], **cs)
    k1 = {"adaptive": batches[names[0]], "colsort": batches[names[1] if len(names) > 1 else "k1_colsort"]}
    for d in dms:
        d.destroy()
    batch = batches[names[0]]
    rows, cols, nnz = batch.rows, batch.cols, batch.nnz
    log(f"[bench] batch of {batch.count}: rows={rows} nnz={nnz} (built in {time.time() - t0:.1f}s)")
    del members

    def time_graph(b, k, X, Y):
        g = b.graph(k, X, Y)
        for _ in range(args.warmup):
            mfp.flush_cache(1 << 30)
            g.launch()
        mfp.synchronize()
        ms = []
        for _ in range(args.steps):
            mfp.flush_cache(1 << 30)
            timer.start()
            g.launch()
            timer.stop()
            ms.append(timer.ms())
        g.destroy()
        return statistics.fmean(ms)

    res = {}
    for k in (1, 4, 16):
        X = mfp.DeviceArray.from_host(mfp.generate_vector(2, cols * k, np.float32))
        Y = mfp.DeviceArray(rows * k, np.float32)
        cands = k1 if k == 1 else {n: batches[n] for n in names}
        timed = {name: time_graph(b, k, X, Y) for name, b in cands.items()}
        mode = min(timed, key=timed.get)
        mean_ms = timed[mode]
        nb = nnz * 8 + (rows + 1) * 4 + k * (cols + rows) * 4
        res[k] = {"ms": round(mean_ms, 5), "GB_per_s": round(nb / (mean_ms * 1e-3) / 1e9, 1),
                  "gflops": round(2 * nnz * k / (mean_ms * 1e-3) / 1e9, 1), "algorithmic_bytes": nb,
                  "plan": {"adaptive": "persistent CSR-Adaptive",
                           "colsort": "column-sorted SCOO" + (", 8192-row slices" if k == 1 else " SpMM"),
                           "mean": f"CSR-vector SpMM, {mean_lanes} lanes",
                           "half": f"CSR-vector SpMM, {mean_lanes // 2} lanes",
                           "binned": "CSR-vector SpMM, lanes per row (2..64)",
                           "stream": "CSR-stream SpMM, row blocks of 512 entries"
                           }.get(mode, f"CSR-vector SpMM, {args.batch_lanes} lanes"),
                  "ms_by_plan": {n: round(v, 5) for n, v in timed.items()}}
        if k == 1:
            # against the FP64 oracle on the GPU over the assembled matrix
            batch = cands[mode]
            batch.spmm(1, X, Y)
            m = mfp.engine.C.c_void_p()
            mfp._lib.call("mfp_batch_matrix", batch.handle, mfp.engine.C.byref(m))
            A = mfp.DeviceMatrix(m.value)
            y64, ad = A.oracle_f64(X)
            res[k]["max_row_err_vs_fp64"] = mfp.residual(Y, y64, ad)["max_row_err"]
            A.handle = None  # owned by the batch
        log(f"[bench] batch k={k}: {mean_ms * 1e3:.1f} us {res[k]['GB_per_s']} GB/s {res[k]['gflops']} GFLOP/s "
            f"({mode}; {', '.join(f'{n} {v * 1e3:.1f} us' for n, v in timed.items())})")
        X.free()
        Y.free()
    value = res[1]["GB_per_s"]
    out = {
        "metric": METRIC, "value": value, "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res[1]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batch512", "description": WORKLOADS["batch512"], "matrices": batch.count,
                   "plans": "per k in batch[k].plan (the fastest of those timed in ms_by_plan)",
                   "rows": rows, "cols": cols, "nnz": nnz, "launch": "one HIP graph replay per step (one persistent SpMM kernel, or the column-sorted SCOO main + finish kernels)",
                   "cache": "Infinity Cache evicted before every step",
                   "measured_copy_GB_per_s": round(copy_gbs, 1)},
        "pct_hbm_roofline": round(100 * value / HBM_PEAK_GBS, 2),
        "batch": res,
        "roofline": {"bound": "hbm", "achieved": value, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(value / HBM_PEAK_GBS, 4), "traffic": None,
                     "basis": "the k = 1 step (one graph replay of the launch), B_alg / step time"},
    }
    for b in {id(b): b for b in batches.values()}.values():
        b.destroy()
    if production:
        return out
    print(json.dumps(out), flush=True)


def tuning_summary(fmt, t):
    """The tuning fields that matter for a format."""
    keys = {"csr_vector": ["lanes"], "csr_adaptive": ["nnz_per_wg", "long_row", "chunk"],
            "ell": ["ell_width"], "hyb": ["hyb_width", "hyb_percent"], "hyb_atomic": ["hyb_width", "hyb_percent"],
            "scoo": ["scoo_slice_size", "scoo_lanes", "scoo_slices", "scoo_pieces", "scoo_threads"]}.get(fmt, [])
    out = {k: t[k] for k in keys}
    if fmt == "scoo":
        out["column_sorted"] = bool(t["flags"] & 4)
    if fmt == "ell" and t["flags"] & 32:
        out["packed_16bit_columns"] = True  # MFP_TUNE_ELL_PACKED (16-bit value plans have a layout of their own)
    return out


def b_alg(rows, cols, nnz, s_v, s_x):
    return nnz * (s_v + 4) + (rows + 1) * 4 + cols * s_x + rows * s_x


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_names(fmt, info):
    """The kernels one SpMV of an FP32/FP64 plan launches, in order."""
    if fmt == "scoo":
        return ["k_scoo_cs", "k_scoo_cs_finish"] if info["scoo_colsort"] else ["k_scoo"]
    if fmt == "csr_adaptive":
        return ["k_csr_adaptive", "k_adaptive_finish"] if info["chunks"] else ["k_csr_adaptive"]
    if fmt == "ell" and info["column_bytes"] == 2:
        return ["k_ell_pk"]
    return {"csr_scalar": ["k_csr_scalar"], "csr_vector": ["k_csr_vector"], "ell": ["k_ell"], "coo": ["k_coo"],
            "hyb": ["k_ell", "k_coo"] if info["coo_elements"] else ["k_ell"],
            "hyb_atomic": ["k_hyb_atomic"]}.get(fmt)


def kernel_bytes(fmt, info, nbytes, rows, s_y, nk):
    """Algorithmic bytes credited to each kernel of one SpMV, and how exact
    that split is.  B_alg = matrix + x reads + y writes."""
    y_bytes = rows * s_y
    if nk == 1:
        return [nbytes], "exact: one kernel does the whole SpMV"
    if fmt == "scoo" and info["scoo_colsort"] and nk == 2:
        return [nbytes - y_bytes, y_bytes], ("main: matrix and x reads; finish: the y writes (a slice with one "
                                             "work item writes its rows of y in the main kernel)")
    if fmt == "csr_adaptive" and nk == 2:
        return [nbytes, 0], "main: everything but the y entries of the chunked long rows, which the finish writes"
    if fmt in ("coo", "hyb_atomic") and nk == 2:
        return [y_bytes, nbytes - y_bytes], "zero fill of y, then the kernel's matrix and x reads"
    if fmt == "hyb" and nk == 2:
        ell_share = 1.0 - info["coo_elements"] / max(info["nnz"], 1)
        m = nbytes - y_bytes - (rows + 1) * 4 - info["cols"] * s_y
        ell = int(ell_share * m) + (rows + 1) * 4 + info["cols"] * s_y + y_bytes
        return [ell, nbytes - ell], "approximate: the matrix bytes split by the ELL / COO entry counts"
    return [nbytes / nk] * nk, "approximate: B_alg split evenly over the kernels"


def rocsparse_row(mfp, A, xd, yd, rp, ci, va, x, nbytes, reps, flush, timer):
    """The "GPU CSR (rocSPARSE)" comparator (the reference's cuSPARSE slot,
    gpu_matrix_multiplier.cu:175-257): rocsparse_spmv's CSR-adaptive
    algorithm on the same device CSR, timed like the candidates (cache
    flush, event pair, median).  Benchmark-only: librocsparse.so is opened at
    run time by app/rocsparse_slot.cpp; None without it."""
    import ctypes as C
    path = os.path.join(ROOT, "matrix_format_performance_amd", "librocsparse_slot.so")
    if not os.path.exists(path) or A.value_type not in ("f32", "f64"):
        return None
    lib = C.CDLL(path)
    lib.mfp_rs_error.restype = C.c_char_p
    if not lib.mfp_rs_available():
        return {"skipped": lib.mfp_rs_error().decode()}
    lib.mfp_rs_plan_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_uint32, C.c_uint32, C.c_uint64,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.mfp_rs_plan_spmv.argtypes = [C.c_void_p, C.c_void_p]
    lib.mfp_rs_plan_destroy.argtypes = [C.c_void_p]
    p = C.c_void_p()
    dt = mfp._lib.MFP_F32 if A.value_type == "f32" else mfp._lib.MFP_F64
    if lib.mfp_rs_plan_create(C.byref(p), dt, A.rows, A.cols, A.nnz, A.row_ptr_ptr, A.columns_ptr, A.values_ptr,
                              xd.ptr, yd.ptr, None) != 0:
        return {"skipped": lib.mfp_rs_error().decode()}
    ms = []
    for _ in range(reps + 1):
        if flush:
            mfp.flush_cache(1 << 30)
        timer.start()
        lib.mfp_rs_plan_spmv(p, None)
        timer.stop()
        ms.append(timer.ms())
    mfp.synchronize()
    lib.mfp_rs_plan_destroy(p)
    import oracle
    y = yd.to_host()
    _, absdot = oracle.spmv_f64(rp, ci, va, x)
    err = oracle.tolerance_check(y, oracle.spmv_seq(rp, ci, va, x), absdot, 1.0)
    med = statistics.median(ms[1:])
    log(f"[bench] rocsparse      {med * 1e3:9.2f} us  {nbytes / (med * 1e-3) / 1e9:8.1f} GB/s  (comparator)")
    return {"ms": round(med, 5), "GB_per_s": round(nbytes / (med * 1e-3) / 1e9, 1),
            "frac_hbm": round(nbytes / (med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "label": "GPU CSR (rocSPARSE)",
            "tuning": {"algorithm": "rocsparse_spmv_alg_csr_adaptive"}, "max_norm_err_vs_cpu_reference": err,
            "note": "benchmark-only comparator, never a selection candidate"}


def time_plan(mfp, plan, xd, yd, reps, flush, timer):
    ms = []
    for _ in range(reps):
        if flush:
            mfp.flush_cache(1 << 30)
        timer.start()
        plan.spmv(xd, yd)
        timer.stop()
        ms.append(timer.ms())
    return ms


def measured_copy_gbs(mfp, timer, nbytes=1 << 30, reps=5):
    """HBM copy rate of a 16-byte streaming copy (read + write bytes / time)."""
    from matrix_format_performance_amd import _lib as L
    src = mfp.DeviceArray(nbytes // 4, np.float32)
    dst = mfp.DeviceArray(nbytes // 4, np.float32)
    src.fill(1.0)
    ms = []
    for _ in range(reps + 1):
        timer.start()
        L.call("mfp_stream_copy", dst.ptr, src.ptr, nbytes, None)
        timer.stop()
        ms.append(timer.ms())
    src.free()
    dst.free()
    return 2 * nbytes / (statistics.median(ms[1:]) * 1e-3) / 1e9


def host_cpu():
    """CPU model, nproc and the threads the CPU baseline uses: the affinity
    mask's size (the analog of std::thread::hardware_concurrency(),
    cpu_matrix_multiplier.cpp:263), capped by OMP_NUM_THREADS where the host
    grants this process that many threads."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    cap = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(cap)) if cap and cap.isdigit() and int(cap) > 0 else affinity
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads": int(cap) if cap and cap.isdigit() else None, "threads": threads}


def cpu_baseline(rp, ci, va, x, nbytes, threads, cpu):
    import oracle
    _, t1 = oracle.spmv_seq_timed(rp, ci, va, x, reps=3)
    times = []
    t_end = time.time() + 8.0
    while time.time() < t_end or len(times) < 5:
        _, t = oracle.spmv_multi_thread(rp, ci, va, x, threads)
        times.append(t)
    tm = statistics.median(times)
    return {
        "value": round(nbytes / tm / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu,
        "sample": (f"full matrix, oracle restatement of cpu_csr_spmv_multi_thread_naive "
                   f"(cpu_matrix_multiplier.cpp:229-308), {threads} threads, median of {len(times)} "
                   f"passes = {tm * 1e3:.3f} ms; single-thread cpu_csr_spmv_single_thread_naive "
                   f"(:20-61) mean of 3 passes = {t1 * 1e3:.3f} ms ({nbytes / t1 / 1e9:.3f} GB/s)"),
    }


def precision_study(mfp, timer, flush, reps, cache):
    """FP32 plain form against BF16 / FP16 value storage (FP32 accumulation
    with compensated row sums, or exact products added in double in the
    column-sorted SCOO) on the 3D 27-point 128^3 stencil and the power-law
    matrix: best time per value type, speed-up over FP32, and the error
    against the GPU FP64 oracle of the stored values (arithmetic) and of the
    FP32 values (arithmetic + quantisation)."""
    out = {}
    for wl in ("stencil3d_128", "powerlaw_1m_normal"):
        rp, ci, va = cache[wl] if wl in cache else mfp.workload(wl, np.float32)
        rows, nnz = rp.size - 1, ci.size
        A = mfp.DeviceMatrix.from_csr(rp, ci, va, rows)
        x = mfp.generate_vector(0, rows, np.float32)
        xd = mfp.DeviceArray.from_host(x)
        yd = mfp.DeviceArray(rows, np.float32)
        y64, ad = A.oracle_f64(xd)
        res = {}
        # "f32_off16": FP32 values in the packed ELL (16-bit column offsets,
        # 6 bytes per slot), reported beside the plain FP32 forms, which the
        # 16-bit value forms' speed-up is measured against
        for vt in ("f32", "f32_off16", "bf16", "f16"):
            M = A if vt.startswith("f32") else A.convert(vt)
            s_v = 4 if vt.startswith("f32") else 2
            nb = b_alg(rows, rows, nnz, s_v, 4)
            best = None
            cands = ((("ell", {"flags": mfp._lib.MFP_TUNE_ELL_PACKED}),) if vt == "f32_off16" else
                     (("csr_vector", {}), ("csr_adaptive", {}), ("ell", {}),
                      ("scoo", {"flags": mfp._lib.MFP_TUNE_SCOO_COLSORT, "scoo_slice_size": 16384,
                                "scoo_threads": 1024})))
            for fmt, kw in cands:
                try:
                    p = mfp.Plan(M, fmt, **kw)
                except mfp.MfpBudgetError:
                    continue
                except mfp.MfpError:
                    if vt != "f32_off16":
                        raise
                    continue  # the packed layout is refused where a 64-row slice spans too many columns
                time_plan(mfp, p, xd, yd, 1, flush, timer)
                ms = statistics.median(time_plan(mfp, p, xd, yd, reps, flush, timer))
                if best is None or ms < best[1]:
                    if best is not None:
                        best[2].destroy()
                    best = (fmt, ms, p)
                else:
                    p.destroy()
            if best is None:
                continue
            fmt, ms, p = best
            p.spmv(xd, yd)
            mfp.synchronize()
            total = mfp.residual(yd, y64, ad)
            if vt.startswith("f32"):
                arith = total
            else:
                ys, ads = M.oracle_f64(xd)
                arith = mfp.residual(yd, ys, ads)
                ys.free()
                ads.free()
            p.destroy()
            res[vt] = {"format": fmt, "ms": round(ms, 5),
                       "effective_GB_per_s": round(nb / (ms * 1e-3) / 1e9, 1),
                       "algorithmic_bytes": nb, "rel_l2_vs_fp64_stored": arith["rel_l2"],
                       "max_row_err_vs_fp64_stored": arith["max_row_err"],
                       "rel_l2_vs_fp64_fp32_values": total["rel_l2"]}
            tpath = os.path.join(ROOT, "profiles", f"traffic_{wl}_{vt}_{fmt}.json")
            if os.path.exists(tpath):
                with open(tpath) as f:
                    hb = json.load(f)["hbm_bytes_per_launch"]
                res[vt]["counter_hbm_bytes"] = hb
                res[vt]["counter_GB_per_s"] = round(hb / (ms * 1e-3) / 1e9, 1)
            if vt == "f32_off16":
                res[vt]["format"] = "ell, 16-bit column offsets"
                res[vt]["speedup_vs_f32"] = round(res["f32"]["ms"] / ms, 3)
            elif vt != "f32":
                res[vt]["speedup_vs_f32"] = round(res["f32"]["ms"] / ms, 3)
                if "f32_off16" in res:
                    res[vt]["speedup_vs_f32_off16"] = round(res["f32_off16"]["ms"] / ms, 3)
                M.destroy()
            log(f"[bench] precision {wl} {vt}: {fmt} {ms * 1e3:.1f} us "
                f"err(stored) {arith['rel_l2']:.2e} err(total) {total['rel_l2']:.2e}")
        for a in (xd, yd, y64, ad):
            a.free()
        A.destroy()
        out[wl] = res
    out["note"] = ("effective_GB_per_s = B_alg / time with B_alg = nnz (s_v + 4) + (rows + 1) 4 + x + y; the "
                   "16-bit sliced ELL and the packed FP32 ELL (f32_off16) store 2-byte column offsets beside "
                   "their values, so their effective rates can exceed the HBM rate; speedup_vs_f32 is against "
                   "the plain FP32 forms (4-byte columns), speedup_vs_f32_off16 against the packed FP32 ELL; counter_GB_per_s (where profiles/ holds the counters) is the HBM traffic "
                   "the FETCH_SIZE / WRITE_SIZE counters measured per launch over the same time")
    return out


def f64_leg(mfp, timer, flush, reps):
    """The irregular workload in FP64: the fastest autotuner candidate,
    timed like the headline steps."""
    rp, ci, va = mfp.workload("powerlaw_1m", np.float64)
    rows, nnz = rp.size - 1, ci.size
    A = mfp.DeviceMatrix.from_csr(rp, ci, va, rows)
    xd = mfp.DeviceArray.from_host(mfp.generate_vector(0, rows, np.float64))
    yd = mfp.DeviceArray(rows, np.float64)
    cands = mfp.autotune(A, xd, yd, reps=reps, flush=flush)
    ok = [c for c in cands if c["status"] == 0 and np.isfinite(c["ms"])]
    best = ok[0]
    p = mfp.Plan(A, best["format"], **best["tuning"])
    time_plan(mfp, p, xd, yd, 1, flush, timer)
    ms = statistics.fmean(time_plan(mfp, p, xd, yd, 20, flush, timer))
    y64, ad = A.oracle_f64(xd)
    p.spmv(xd, yd)
    err = mfp.residual(yd, y64, ad)["max_row_err"]
    nb = b_alg(rows, rows, nnz, 8, 8)
    out = {"workload": "powerlaw_1m", "values": "f64", "format": best["format"],
           "tuning": tuning_summary(best["format"], best["tuning"]), "ms": round(ms, 5),
           "GB_per_s": round(nb / (ms * 1e-3) / 1e9, 1), "frac_hbm": round(nb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
           "algorithmic_bytes": nb, "max_row_err_vs_fp64_gpu_oracle": err, "candidates": len(cands)}
    log(f"[bench] f64 powerlaw_1m: {best['format']} {ms * 1e3:.1f} us {out['GB_per_s']} GB/s")
    for o in (p, A):
        o.destroy()
    for a in (xd, yd, y64, ad):
        a.free()
    return out


def held_out_regret(mfp):
    """The shipped selector on the held-out fifth (480) of the committed
    2,400-matrix sweep, BF16 inference on the matrix cores: the regret of
    its picks against each matrix's exhaustive best."""
    from matrix_format_performance_amd import selector as S
    import gzip
    recs = []
    for name in ("selector_c5_sweep.jsonl.gz", "selector_c5_sweep_big.jsonl.gz"):
        path = os.path.join(ROOT, "results", name)
        if not os.path.exists(path):
            return None
        byidx = {}
        with gzip.open(path, "rt") as f:
            for line in f:
                r = json.loads(line)
                byidx[r["index"]] = r
        recs += [byidx[k] for k in sorted(byidx)]
    te = [r for i, r in enumerate(recs) if i % 5 == 4]
    sl = S.Selector.default()
    probs = sl.net.predict(np.array([r["features"] for r in te], np.float32), bf16=True)
    ev = S.evaluate_with_rules(te, probs, sl.rules)
    sl.net.destroy()
    ev = {k: round(v, 4) for k, v in ev.items()}
    ev.update({"matrices": len(te), "source": "results/selector_c5_sweep*.jsonl.gz, every fifth record held out; "
               "BF16 inference on the matrix cores, tuning by the rules"})
    return ev


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="powerlaw_1m", choices=sorted(WORKLOADS))
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--format", default="auto")
    ap.add_argument("--select-reps", type=int, default=5)
    ap.add_argument("--no-flush", action="store_true", help="warm-cache timing (labelled)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0 = the affinity mask, capped by OMP_NUM_THREADS)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=None, help="json with FETCH/WRITE-derived bytes per launch")
    ap.add_argument("--no-precision", action="store_true", help="skip the BF16 / FP16 study")
    ap.add_argument("--no-batch", action="store_true",
                    help="default run: skip the 512-matrix batch section (k = 1, 4, 16)")
    ap.add_argument("--batch-lanes", type=int, default=-1,
                    help="batch512, k > 1: -1 = time the mean-based lane count (and half of it), per-row "
                         "lanes, the column-sorted SpMM and the CSR-stream SpMM and keep the fastest; 0 = "
                         "per-row lanes; 1 = column-sorted SpMM; 3 = CSR-stream SpMM; N = N lanes for every row")
    ap.add_argument("--values", default=None, choices=["bf16", "f16"],
                    help="store the matrix values in 16 bits (FP32 vectors and accumulation)")
    ap.add_argument("--bandwidth", type=float, default=1.0,
                    help="power-law only: column window as a fraction of cols (study knob)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU row split is a design, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import matrix_format_performance_amd as mfp

    dtype = np.float32 if args.dtype == "f32" else np.float64
    s = np.dtype(dtype).itemsize
    mfp.init(0)
    info = mfp.device_info()
    log(f"[bench] device {info['name']} {info['arch']} CUs={info['cu_count']}")
    t0 = time.time()
    if args.workload == "batch512":
        run_batch(mfp, args, mfp.Timer())
        return
    if args.workload == "powerlaw_1m" and args.bandwidth < 1.0:
        rp, ci, va = mfp.generate(mfp.GenSpec("powerlaw", values="uniform", rows=1 << 20, nnz_lo=1,
                                              nnz_hi=50000, alpha=2.0, seed=0,
                                              bandwidth=args.bandwidth), dtype)
    else:
        rp, ci, va = mfp.workload(args.workload, dtype)
    rows, nnz = rp.size - 1, ci.size
    cols = rows
    x = mfp.generate_vector(0, cols, dtype)  # x ~ U[-1, 1) from seed 0
    nbytes = b_alg(rows, cols, nnz, 2 if args.values else s, s)
    log(f"[bench] workload {args.workload}: rows={rows} nnz={nnz} B_alg={nbytes / 1e6:.1f} MB "
        f"(generated in {time.time() - t0:.1f}s)")

    A = mfp.DeviceMatrix.from_csr(rp, ci, va, cols)
    if args.values:
        if args.dtype != "f32":
            sys.exit("--values needs --dtype f32 (16-bit values use FP32 vectors)")
        A32, A = A, A.convert(args.values)
        A32.destroy()
    xd = mfp.DeviceArray.from_host(x)
    yd = mfp.DeviceArray(rows, dtype)
    timer = mfp.Timer()
    flush = not args.no_flush
    copy_gbs = measured_copy_gbs(mfp, timer)
    log(f"[bench] measured 16-byte copy: {copy_gbs:.1f} GB/s")

    # exhaustive selection: every autotuner candidate of every format, timed
    # the same way as the measured steps (flush, event pair, median)
    fmts = list(mfp.FORMATS) if args.format == "auto" else [args.format]
    t_sel = time.time()
    cands = mfp.autotune(A, xd, yd, formats=fmts, reps=args.select_reps, flush=flush)
    t_sel = time.time() - t_sel
    table = {}
    for c in cands:  # fastest first
        f = c["format"]
        if c["status"] != 0 or not np.isfinite(c["ms"]):
            table.setdefault(f, {"skipped": f"status {c['status']} (layout above the memory budget)"
                                 if c["status"] == 3 else f"status {c['status']}"})
            continue
        if "ms" in table.get(f, {}):
            continue
        med = c["ms"]
        table[f] = {"ms": round(med, 5), "GB_per_s": round(nbytes / (med * 1e-3) / 1e9, 1),
                    "frac_hbm": round(nbytes / (med * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    "tuning": tuning_summary(f, c["tuning"])}
        log(f"[bench] {f:13s} {med * 1e3:9.2f} us  {table[f]['GB_per_s']:8.1f} GB/s  {table[f]['tuning']}")
    rs = rocsparse_row(mfp, A, xd, yd, rp, ci, va, x, nbytes, args.select_reps, flush, timer)
    if rs is not None:
        table["rocsparse"] = rs
    ok = [c for c in cands if c["status"] == 0 and np.isfinite(c["ms"])]
    best_c = ok[0]
    best = best_c["format"]
    plan = mfp.Plan(A, best, **best_c["tuning"])
    log(f"[bench] exhaustive selection over {len(cands)} candidates in {t_sel:.1f}s: {best} "
        f"{tuning_summary(best, best_c['tuning'])}")

    # the learned pick (features + network + tuning rules, nothing timed)
    # against the exhaustive best
    from matrix_format_performance_amd import selector as sel_mod
    selector_out = None
    if os.path.exists(sel_mod.DEFAULT_WEIGHTS):
        t_pick = time.time()
        sl = sel_mod.Selector.default()
        pfmt, ptun = sl.pick(A, row_ptr=rp)
        t_pick = time.time() - t_pick
        pc = next((c for c in ok if c["format"] == pfmt and c["tuning"] == ptun), None)
        if pc is None:
            pp = mfp.Plan(A, pfmt, **ptun)
            time_plan(mfp, pp, xd, yd, 1, flush, timer)
            pms = statistics.median(time_plan(mfp, pp, xd, yd, args.select_reps, flush, timer))
            pp.destroy()
        else:
            pms = pc["ms"]
        selector_out = {"pick": pfmt, "tuning": tuning_summary(pfmt, ptun), "ms": round(pms, 5),
                        "slowdown_vs_exhaustive": round(pms / best_c["ms"], 4),
                        "pick_seconds": round(t_pick, 4), "exhaustive_seconds": round(t_sel, 3)}
        sl.net.destroy()
        log(f"[bench] learned pick {pfmt} {selector_out['tuning']}: {pms * 1e3:.1f} us "
            f"({selector_out['slowdown_vs_exhaustive']}x the exhaustive best)")
        selector_out["held_out"] = held_out_regret(mfp)
        if selector_out["held_out"]:
            log(f"[bench] selector on {selector_out['held_out']['matrices']} held-out matrices: mean regret "
                f"{selector_out['held_out']['mean_regret']}, p90 {selector_out['held_out']['p90_regret']}")

    # correctness of the measured format on this very workload
    time_plan(mfp, plan, xd, yd, 1, flush, timer)
    mfp.synchronize()
    import oracle
    y = yd.to_host()
    va_stored = va
    if args.values == "bf16":
        from oracle.selector import bf16_round
        va_stored = bf16_round(va)
    elif args.values == "f16":
        va_stored = va.astype(np.float16).astype(np.float32)
    y_ref = oracle.spmv_seq(rp, ci, va_stored, x)
    _, absdot = oracle.spmv_f64(rp, ci, va_stored, x)
    err = oracle.tolerance_check(y, y_ref, absdot, 1.0)

    for _ in range(args.warmup):
        time_plan(mfp, plan, xd, yd, 1, flush, timer)
    mfp.synchronize()
    t_wall = time.time()
    ms = time_plan(mfp, plan, xd, yd, args.steps, flush, timer)
    mfp.synchronize()
    t_wall = time.time() - t_wall
    mean_ms = statistics.fmean(ms)
    value = nbytes / (mean_ms * 1e-3) / 1e9

    # roofline of the dominant kernel: the same flushed steps again, each with
    # an event after every kernel of the SpMV (mfp_plan_kernel_times)
    per = []
    for _ in range(args.steps):
        if flush:
            mfp.flush_cache(1 << 30)
        per.append(plan.kernel_times(xd, yd))
    nk = min(len(k) for k in per)
    kernel_ms = [statistics.fmean(k[i] for k in per) for i in range(nk)]
    dom = max(range(nk), key=lambda i: kernel_ms[i])
    # each kernel is credited with its share of B_alg (kernel_bytes)
    pinfo = plan.info()
    shares, share_note = kernel_bytes(best, pinfo, nbytes, rows, s, nk)
    if args.values:
        share_note = "approximate: 16-bit value plans time each kernel from the previous event; " + share_note
    dom_bytes = shares[dom]
    dom_names = None if args.values else kernel_names(best, pinfo)
    achieved = dom_bytes / (kernel_ms[dom] * 1e-3) / 1e9

    traffic = None
    vtag = args.values or args.dtype
    if best == "ell" and not args.values and pinfo["column_bytes"] == 2:
        vtag += "_off16"  # the packed FP32 ELL's own counters
    tpath = args.traffic or os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{vtag}_{best}.json")
    if not os.path.exists(tpath) and best == "hyb" and dom_names == ["k_ell"]:
        # a hybrid whose COO part is empty runs the ELL kernel on the ELL layout
        tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{vtag}_ell.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "description": WORKLOADS[args.workload],
            "rows": rows, "cols": cols, "nnz": nnz,
            "values": args.values or args.dtype,
            "format": best,
            "tuning": tuning_summary(best, best_c["tuning"]),
            "selection": f"exhaustive autotune, {len(cands)} candidates, {t_sel:.1f} s",
            "algorithmic_bytes": nbytes,
            "cache": "warm (no flush)" if not flush else "Infinity Cache evicted before every step",
            "median_ms": round(statistics.median(ms), 5),
            "min_ms": round(min(ms), 5),
            "wall_s_timed_region_incl_flushes": round(t_wall, 4),
            "max_norm_err_vs_cpu_reference": err,
            "gflops": round(2 * nnz / (mean_ms * 1e-3) / 1e9, 2),
            "measured_copy_GB_per_s": round(copy_gbs, 1),
            "frac_of_measured_copy": round(value / copy_gbs, 4),
        },
        "pct_hbm_roofline": round(100 * value / HBM_PEAK_GBS, 2),
        "learned_pick": selector_out,
        "formats": table,
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": dom_names[dom] if dom_names and len(dom_names) == nk else f"kernel {dom + 1} of {nk}",
            "ms_per_launch": round(kernel_ms[dom], 5),
            "algorithmic_bytes_per_launch": int(dom_bytes),
            "kernel_ms": [round(t, 5) for t in kernel_ms],
            "kernel_bytes": [int(b) for b in shares],
            "bytes_split": share_note,
            "step_achieved": round(value, 2),
            "step_frac": round(value / HBM_PEAK_GBS, 4),
        },
    }
    # host-to-host rate: upload CSR and x, build the plan, one SpMV, download y
    plan_kw = best_c["tuning"]
    mfp.synchronize()
    t_h = time.perf_counter()
    A2 = mfp.DeviceMatrix.from_csr(rp, ci, va, cols)
    if args.values:
        A2f, A2 = A2, A2.convert(args.values)
        A2f.destroy()
    x2 = mfp.DeviceArray.from_host(x)
    y2 = mfp.DeviceArray(rows, dtype)
    p2 = mfp.Plan(A2, best, **plan_kw)
    p2.spmv(x2, y2)
    y_host = y2.to_host()
    t_h = time.perf_counter() - t_h
    assert y_host.shape == (rows,)
    out["config"]["host_to_host_GB_per_s"] = round(nbytes / t_h / 1e9, 2)
    out["config"]["host_to_host_ms"] = round(t_h * 1e3, 3)
    for o in (p2, A2, x2, y2):
        o.destroy() if hasattr(o, "destroy") else o.free()
    plan.destroy()

    if args.dtype == "f32" and not args.values and not args.no_precision:
        A.destroy()
        out["precision"] = precision_study(mfp, timer, flush, args.select_reps, {})
        if args.workload == "powerlaw_1m" and args.bandwidth >= 1.0:
            out["f64"] = f64_leg(mfp, timer, flush, args.select_reps)
            if not args.no_batch:
                # the 512-matrix batch (k = 1, 4 and 16) beside the headline
                b = run_batch(mfp, args, timer, production=True)
                out["batch512"] = {"config": b["config"], "batch": b["batch"]}
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = host_cpu()
        threads = args.cpu_threads or cpu["threads"]
        out["cpu_baseline"] = cpu_baseline(rp, ci, va, x, nbytes, threads, cpu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
