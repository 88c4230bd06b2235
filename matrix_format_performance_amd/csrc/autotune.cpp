// Autotuner: candidate (format, tuning) sets for a matrix and their timing.
//
// Swept parameters: lanes per row (CSR-vector, 2..64), CSR-Adaptive block
// size, ELL width (longest row) and the packed FP32 ELL layout, the ELL/COO
// split of the hybrids (row-length quantiles), the SCOO slice size (lane
// count filling a 32 KB LDS accumulator) and column-sorted SCOO slices of
// 4096..16384 rows (pieces per slice and workgroup size).  Every candidate is built as a plan and timed with an event
// pair (median of reps, each after an Infinity-Cache flush when asked).
#include "../../include/mfp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace {

uint32_t prev_pow2(uint32_t v)
{
  uint32_t p = 1;
  while ((p << 1) && (p << 1) <= v) p <<= 1;
  return v ? p : 0;
}

void push(std::vector<mfp_candidate> &out, int fmt, const mfp_tuning &t)
{
  mfp_candidate c;
  std::memset(&c, 0, sizeof(c));
  c.format = fmt;
  c.tuning = t;
  c.ms = std::numeric_limits<double>::infinity();
  c.status = -1;
  out.push_back(c);
}

} // namespace

extern "C" {

int mfp_autotune_candidates(mfp_matrix m, uint32_t format_mask, mfp_candidate *out, uint32_t cap,
                            uint32_t *n)
{
  if (!m || !n) return MFP_ERR_ARG;
  mfp_matrix_view v;
  int st = mfp_matrix_view_get(m, &v);
  if (st != MFP_OK) return st;
  const bool lp = v.dtype == MFP_BF16 || v.dtype == MFP_F16;
  // row-length quantiles from the row pointers
  std::vector<uint32_t> rp((size_t)v.rows + 1), len(v.rows);
  st = mfp_memcpy_d2h(rp.data(), v.row_ptr, rp.size() * 4);
  if (st != MFP_OK) return st;
  for (uint32_t r = 0; r < v.rows; r++) len[r] = rp[r + 1] - rp[r];
  auto quant = [&](double q) -> uint32_t {
    if (len.empty()) return 0;
    std::vector<uint32_t> tmp(len);
    const size_t k = std::min(tmp.size() - 1, (size_t)(q * (double)(tmp.size() - 1)));
    std::nth_element(tmp.begin(), tmp.begin() + k, tmp.end());
    return tmp[k];
  };
  const uint32_t avg = v.rows ? (uint32_t)((v.nnz + v.rows - 1) / v.rows) : 1;

  std::vector<mfp_candidate> c;
  auto want = [&](int f) { return (format_mask >> f) & 1u; };
  for (int f = 0; f < MFP_FORMAT_COUNT; f++) {
    if (!want(f)) continue;
    if (lp && f != MFP_CSR_VECTOR && f != MFP_ELL && f != MFP_CSR_ADAPTIVE && f != MFP_SCOO) continue;
    mfp_tuning t;
    st = mfp_tuning_default(m, f, &t);
    if (st != MFP_OK) return st;
    const mfp_tuning def = t;
    mfp_tuning tagged = def;
    tagged.flags |= MFP_TUNE_IS_DEFAULT;
    switch (f) {
    case MFP_CSR_VECTOR:
      for (uint32_t l = 2; l <= 64; l <<= 1) {
        t.lanes = l;
        push(c, f, l == def.lanes ? tagged : t);
      }
      break;
    case MFP_CSR_ADAPTIVE: {
      if (lp) { // the binned 16-bit form has no block size
        push(c, f, tagged);
        break;
      }
      const uint32_t base = t.nnz_per_wg;
      for (uint32_t w : {base / 2, base, base * 2}) {
        t.nnz_per_wg = w;
        push(c, f, w == base ? tagged : t);
      }
      break;
    }
    case MFP_SCOO: {
      push(c, f, tagged);
      const uint32_t acc = 4096u; // 32 KB of double accumulators
      for (uint32_t ss : {256u, 512u, 1024u, 2048u, 4096u}) {
        if (lp) break; // 16-bit values: the column-sorted form only
        if (v.rows && ss > std::max<uint32_t>(256, v.rows)) break;
        t.scoo_slice_size = ss;
        t.scoo_lanes = std::max<uint32_t>(1, std::min<uint32_t>(32, prev_pow2(acc / ss)));
        t.scoo_slices = v.rows ? (v.rows + ss - 1) / ss : 1;
        if (ss == def.scoo_slice_size && t.scoo_lanes == def.scoo_lanes && t.scoo_slices == def.scoo_slices)
          continue;
        push(c, f, t);
      }
      // column-sorted slices (8-byte LDS accumulators): pieces chosen by
      // the plan, or 2 and 4 pieces per slice; a 16384-row slice fills
      // 128 KB of LDS (one workgroup per CU), so it also tries 1024 threads
      for (uint32_t R : {4096u, 8192u, 16384u}) {
        if (v.rows && R > 4096 && R / 2 >= v.rows) break;
        for (uint32_t P : {0u, 2u, 4u})
          for (uint32_t th : {512u, 1024u}) {
            if (th == 1024 && R < 16384) continue;
            mfp_tuning tc = def;
            tc.flags = MFP_TUNE_SCOO_COLSORT;
            tc.scoo_slice_size = R;
            tc.scoo_pieces = P;
            tc.scoo_threads = th;
            if (lp && R == def.scoo_slice_size && P == 0 && th == def.scoo_threads) continue; // the default, above
            push(c, f, tc);
          }
      }
      break;
    }
    case MFP_HYB:
    case MFP_HYB_ATOMIC: {
      push(c, f, tagged); // width from the hyb_percent rule
      std::vector<uint32_t> ws = {avg, quant(0.75), quant(0.9), quant(0.99)};
      std::sort(ws.begin(), ws.end());
      ws.erase(std::unique(ws.begin(), ws.end()), ws.end());
      for (uint32_t w : ws) {
        if (w == 0) continue;
        t.hyb_width = w;
        push(c, f, t);
      }
      break;
    }
    case MFP_ELL:
      push(c, f, tagged);
      if (v.dtype == MFP_F32 || v.dtype == MFP_F64) { // the packed layout (its plan refuses slices whose columns span 65,536 or more)
        mfp_tuning tp = def;
        tp.flags |= MFP_TUNE_ELL_PACKED;
        push(c, f, tp);
      }
      break;
    default: push(c, f, tagged); break;
    }
  }
  *n = (uint32_t)c.size();
  if (out) std::memcpy(out, c.data(), std::min<size_t>(cap, c.size()) * sizeof(mfp_candidate));
  return MFP_OK;
}

int mfp_autotune_run(mfp_matrix m, const void *x, void *y, mfp_candidate *cands, uint32_t n,
                     uint32_t reps, int flush)
{
  if (!m || !x || !y || (!cands && n)) return MFP_ERR_ARG;
  if (reps == 0) reps = 1;
  mfp_timer tm = nullptr;
  int st = mfp_timer_create(&tm);
  if (st != MFP_OK) return st;
  for (uint32_t i = 0; i < n; i++) {
    mfp_candidate &c = cands[i];
    mfp_plan p = nullptr;
    c.status = mfp_plan_create(&p, m, c.format, &c.tuning, nullptr);
    c.ms = std::numeric_limits<double>::infinity();
    if (c.status != MFP_OK) continue;
    std::vector<double> t;
    for (uint32_t r = 0; r <= reps && c.status == MFP_OK; r++) { // first run warms the code path
      if (flush) c.status = mfp_flush_cache(1ull << 30, nullptr);
      if (c.status == MFP_OK) c.status = mfp_timer_start(tm, nullptr);
      if (c.status == MFP_OK) c.status = mfp_plan_spmv(p, x, y, nullptr);
      if (c.status == MFP_OK) c.status = mfp_timer_stop(tm, nullptr);
      float ms = 0;
      if (c.status == MFP_OK) c.status = mfp_timer_elapsed_ms(tm, &ms);
      if (r > 0) t.push_back(ms);
    }
    mfp_plan_destroy(p);
    if (c.status == MFP_OK && !t.empty()) {
      std::nth_element(t.begin(), t.begin() + t.size() / 2, t.end());
      c.ms = t[t.size() / 2];
    }
  }
  mfp_timer_destroy(tm);
  return MFP_OK;
}

} // extern "C"
