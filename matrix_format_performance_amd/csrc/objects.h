// Opaque C-ABI objects (matrix, plan), shared by plan.hip and batch.hip.
#pragma once

#include "kernels.h"
#include "mfp_internal.h"

struct mfp_matrix_s {
  int dtype = MFP_F32;
  uint32_t rows = 0, cols = 0;
  uint64_t nnz = 0;
  uint32_t *row_ptr = nullptr, *columns = nullptr;
  void *values = nullptr;
  bool stats_valid = false;
  mfp_row_stats stats{};
};

struct mfp_plan_s {
  int format = 0, dtype = 0;
  mfp_matrix m = nullptr;
  mfp_tuning t{};
  uint32_t ell_w = 0;
  void *ell_val = nullptr;
  uint32_t *ell_col = nullptr;
  uint64_t ncoo = 0;
  void *coo_val = nullptr;
  uint32_t *coo_col = nullptr, *coo_row = nullptr;
  uint32_t slices = 0, ss = 0, L = 0, scoo_threads = 512;
  uint32_t *sp = nullptr;
  mfp::adaptive_layout ad;
  // CSR-Adaptive plans over f32 / f64 values: row blocks of STREAM_W entries
  // as (first row, first entry) pairs for the CSR-stream SpMM (mfp_plan_spmm
  // with k > 1)
  uint32_t *stream_meta = nullptr;
  uint32_t nstream = 0;
  uint64_t device_bytes = 0;
  // 16-bit-value ELL: columns as u16 offsets from base[row] when off16
  bool off16 = false;
  bool lp_block = false; // 16-bit offsets from a base per 64-row block
  void *lp_cols = nullptr;
  uint32_t *lp_base = nullptr;
  // FP32 / FP64 ELL with MFP_TUNE_ELL_PACKED: values and u16 offsets in 64-row
  // slices, pk_w4 words of 4 slots per row, offsets from pk_base[row / 64]
  bool ell_packed = false;
  uint32_t pk_w4 = 0;
  void *pk_val = nullptr;
  uint16_t *pk_off = nullptr;
  uint32_t *pk_base = nullptr;
  // column-sorted SCOO (MFP_TUNE_SCOO_COLSORT)
  bool colsort = false;
  mfp::scoo_cs_layout cs;
  // binned CSR over 16-bit values
  mfp::lp_binned_layout lpb;
  // MFP_TUNE_STATIC_PARTIALS: no buffer of partial results of its own
  bool static_partials = false;
};

