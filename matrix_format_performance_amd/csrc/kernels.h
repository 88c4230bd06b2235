// Kernel launchers (spmv.hip) shared by plans and the reference entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mfp_internal.h"

namespace mfp {

struct adaptive_args {
  uint32_t rows = 0;
  const uint32_t *rp = nullptr, *ci = nullptr;
  const uint32_t *blocks = nullptr; // nblocks + 1 row boundaries
  uint32_t nblocks = 0;
  uint32_t nnz_per_wg = 0;
  uint32_t long_row = 0;
  const uint4 *chunks = nullptr; // {row, piece, first item, pieces}
  uint32_t nchunks = 0;
  uint32_t chunk = 0;
  void *partials = nullptr; // nchunks values
  bool persistent = false;  // fixed grid walking the items (batched launches)
};

// CSR-Adaptive defaults: a row block holds up to this many nonzeros, which a
// 256-lane workgroup reads as two 16-byte vectors per lane.
inline uint32_t default_nnz_per_wg(int dtype) { return dtype == 1 /*MFP_F64*/ ? 1020u : 2040u; }
constexpr uint32_t DEFAULT_LONG_ROW = 4096, DEFAULT_CHUNK = 2048;
uint32_t adaptive_vectors_per_lane(int dtype, uint32_t nnz_per_wg);

int launch_csr_scalar(int dtype, uint32_t rows, const uint32_t *rp, const uint32_t *ci,
                      const void *a, const void *x, void *y, hipStream_t s);
int launch_csr_vector(int dtype, uint32_t lanes, uint32_t rows, const uint32_t *rp,
                      const uint32_t *ci, const void *a, const void *x, void *y, hipStream_t s);
int launch_csr_adaptive(int dtype, const adaptive_args &g, const void *a, const void *x, void *y,
                        hipStream_t s);
// FP32 / FP64 ELL in the 64-row sliced layout with 16-bit offsets (spmv.hip
// k_ell_pk); stride = rows rounded up to 64, base[stride / 64] = each block's
// first column
int launch_ell_pk(int dtype, uint32_t rows, uint32_t w, uint32_t w4, const uint32_t *base, const void *val,
                  const uint16_t *off, const void *x, void *y, hipStream_t s);
int ell_pk_build(int dtype, uint32_t rows, uint32_t stride, const uint32_t *rp, const uint32_t *ci, const void *a,
                 uint32_t w4, const uint32_t *base, void *val, uint16_t *off, hipStream_t s);
int launch_ell(int dtype, uint32_t rows, uint32_t w, const uint32_t *ec, const void *ev,
               const void *x, void *y, hipStream_t s);
int launch_coo_add(int dtype, uint64_t n, const uint32_t *cr, const uint32_t *cc, const void *cv,
                   const void *x, void *y, hipStream_t s);
int launch_vec_add(int dtype, uint32_t n, const void *t, void *y, hipStream_t s);
int launch_hyb_atomic(int dtype, uint32_t rows, uint32_t w, const uint32_t *ec, const void *ev,
                      uint64_t ncoo, const uint32_t *cr, const uint32_t *cc, const void *cv,
                      const void *x, void *y, hipStream_t s);
int launch_scoo_cs(int dtype, uint32_t rows, const scoo_cs_layout &l, uint32_t threads, const void *x, void *y,
                   hipStream_t s);
int launch_scoo_cs_spmm(int dtype, uint32_t k, uint32_t rows, const scoo_cs_layout &l, uint32_t threads,
                        const void *X, void *Y, hipStream_t s);
int launch_scoo(int dtype, uint32_t rows, uint32_t ss, uint32_t L, uint32_t threads,
                const uint32_t *cr, const uint32_t *cc, const void *cv, const uint32_t *sp,
                const void *x, void *y, hipStream_t s);

// Row blocks + long-row chunk list for CSR-Adaptive (plan.hip).  The arrays
// are owned by the caller afterwards (free with dfree).
struct adaptive_layout {
  uint32_t *blocks = nullptr;
  uint32_t nblocks = 0;
  uint4 *chunks = nullptr;
  uint32_t nchunks = 0;
  void *partials = nullptr;
};
// own_partials = false: no chunk-sum buffer (the plan runs with the static
// scratch of mfp_plan_spmv_noalloc)
int build_adaptive_layout(int dtype, uint32_t rows, const uint32_t *rp, uint32_t nnz_per_wg,
                          uint32_t long_row, uint32_t chunk, adaptive_layout *out, hipStream_t s,
                          bool own_partials = true);
void free_adaptive_layout(adaptive_layout *l);

int launch_spmm_vector(int dtype, uint32_t k, uint32_t lanes, uint32_t rows, const uint32_t *rp,
                       const uint32_t *ci, const void *a, const void *X, void *Y, bool persistent,
                       hipStream_t s);
// CSR-stream SpMM over row blocks of STREAM_W entries (fill_row_blocks, batch.hip).
// meta holds nblocks + 1 (first row, first entry) pairs from stream_meta_device.
constexpr uint32_t STREAM_W = 512;
int stream_meta_device(uint32_t rows, const uint32_t *rp, uint32_t **meta, uint32_t *nblocks, hipStream_t s);
int launch_spmm_stream(int dtype, uint32_t k, uint32_t cols, uint32_t nblocks, const uint32_t *meta,
                       const uint32_t *rp, const uint32_t *ci, const void *a, const void *X, void *Y, hipStream_t s);
int offset_copy_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint32_t add, hipStream_t s);

int csr_from_triplets_device(int dtype, uint32_t rows, uint32_t cols, uint64_t nnz,
                             const uint32_t *tr_rows, const uint32_t *tr_cols,
                             const double *tr_vals, uint32_t *row_ptr, uint32_t *columns,
                             void *values, hipStream_t s);

} // namespace mfp
