/*
 * libmfp -- MI355X (gfx950) SpMV engine, C ABI.
 *
 * Plain C types only: sizes, host pointers, device pointers (void*), opaque
 * handles and an int status.  No exceptions cross this boundary.  Every
 * enqueue-style call takes an explicit stream (NULL = the null stream) and
 * performs no host synchronisation, so it can be captured in a HIP graph.
 *
 * Citations name the reference interface each entry point replaces, as
 * path:line relative to the reference repository root.
 */
#ifndef MFP_H
#define MFP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  MFP_OK = 0,
  MFP_ERR_HIP = 1,     /* a HIP runtime call failed (text: mfp_last_error) */
  MFP_ERR_ARG = 2,     /* invalid argument / shape */
  MFP_ERR_BUDGET = 3,  /* the layout would exceed the memory budget */
  MFP_ERR_STORAGE = 4, /* non-general storage (matrix_converter.cpp:19-20) */
  MFP_ERR_NODEV = 5,   /* no usable gfx950 device */
  MFP_ERR_STATE = 6    /* call made in the wrong state */
};

/* ---- value types ------------------------------------------------------- */
enum { MFP_F32 = 0, MFP_F64 = 1, MFP_BF16 = 2, MFP_F16 = 3 };

/* ---- formats / kernels ------------------------------------------------- */
enum {
  MFP_CSR_SCALAR = 0,   /* gpu_matrix_multiplier.cu:22-43   "GPU CSR"             */
  MFP_CSR_VECTOR = 1,   /* gpu_matrix_multiplier.cu:331-362 "GPU CSR (vector)"    */
  MFP_CSR_ADAPTIVE = 2, /* csr_adaptive_spmv.cu:26-178      "GPU CSR-Adaptive"    */
  MFP_ELL = 3,          /* gpu_matrix_multiplier.cu:450-471 "GPU ELL"             */
  MFP_COO = 4,          /* gpu_matrix_multiplier.cu:622-638 "GPU COO"             */
  MFP_SCOO = 5,         /* scoo_spmv.cu:46-129              "GPU SCOO"            */
  MFP_HYB = 6,          /* gpu_matrix_multiplier.cu:866-967 "GPU Hybrid (TODO)"   */
  MFP_HYB_ATOMIC = 7,   /* gpu_matrix_multiplier.cu:969-1096 "GPU Hybrid (atomic)" */
  MFP_FORMAT_COUNT = 8
};

typedef void *mfp_stream; /* hipStream_t */

/* ---- runtime ----------------------------------------------------------- */
typedef struct {
  int device;
  int cu_count;           /* multiProcessorCount (sm_count at main.cpp:418) */
  int xcd_count;
  uint64_t lds_per_block; /* sharedMemPerBlock (main.cpp:419) */
  uint64_t free_mem, total_mem;
  char name[256];
  char arch[64];
} mfp_device_info;

int mfp_init(int device);
int mfp_device_query(mfp_device_info *out);
const char *mfp_status_string(int status);
const char *mfp_last_error(void);
int mfp_synchronize(void);
int mfp_malloc(void **dev_ptr, size_t bytes);
int mfp_free(void *dev_ptr);
int mfp_memcpy_h2d(void *dst, const void *src, size_t bytes);
int mfp_memcpy_d2h(void *dst, const void *src, size_t bytes);
int mfp_memcpy_d2d(void *dst, const void *src, size_t bytes, mfp_stream s);
/* dst[i] = value (value given as double, stored as dtype) */
int mfp_fill(int dtype, void *dst, uint64_t n, double value, mfp_stream s);
/* Read `bytes` of scratch so the 256 MB Infinity Cache holds no matrix data
 * before a timed run (pass >= 512 MB).  The scratch is written once when it
 * is allocated; the flush itself only reads, so it leaves no dirty lines for
 * the timed kernel to write back. */
int mfp_flush_cache(uint64_t bytes, mfp_stream s);
/* dst = src with 16-byte nontemporal loads: the measured-copy reference for
 * HBM bandwidth and for scaling the FETCH_SIZE / WRITE_SIZE counters. */
int mfp_stream_copy(void *dst, const void *src, uint64_t bytes, mfp_stream s);

/* Event pair timing on stream s (hipEventElapsedTime, milliseconds). */
typedef struct mfp_timer_s *mfp_timer;
int mfp_timer_create(mfp_timer *t);
int mfp_timer_start(mfp_timer t, mfp_stream s);
int mfp_timer_stop(mfp_timer t, mfp_stream s);
int mfp_timer_elapsed_ms(mfp_timer t, float *ms); /* waits for the stop event */
int mfp_timer_destroy(mfp_timer t);

/* ---- grow-only device buffer: resizable_gpu_memory<T>
 *      (gpu/resizable_gpu_memory.h:13-44, .cpp:1-20) ------------------------ */
typedef struct mfp_buffer_s *mfp_buffer;
int mfp_buffer_create(mfp_buffer *b);
int mfp_buffer_resize(mfp_buffer b, size_t bytes); /* grows only; contents lost */
int mfp_buffer_clear(mfp_buffer b);
void *mfp_buffer_ptr(mfp_buffer b);
size_t mfp_buffer_size(mfp_buffer b);
int mfp_buffer_destroy(mfp_buffer b);

/* ---- device-resident matrix (CSR, u32 indices) --------------------------- */
typedef struct mfp_matrix_s *mfp_matrix;
typedef struct {
  int dtype;
  uint32_t rows, cols;
  uint64_t nnz;
  const uint32_t *row_ptr; /* device */
  const uint32_t *columns; /* device */
  const void *values;      /* device */
} mfp_matrix_view;

/* Upload a host CSR (copies). */
int mfp_matrix_from_host_csr(mfp_matrix *out, int dtype, uint32_t rows, uint32_t cols,
                             uint64_t nnz, const uint32_t *row_ptr,
                             const uint32_t *columns, const void *values);
/* Device converter: triplets (host or device arrays; values double) -> CSR on
 * the device.  Replaces csr_matrix_class<T>(matrix_class)
 * (common/matrix_converter.cpp:15-87); ties keep input order. */
int mfp_matrix_from_triplets(mfp_matrix *out, int dtype, uint32_t rows, uint32_t cols,
                             uint64_t nnz, const uint32_t *tr_rows,
                             const uint32_t *tr_cols, const double *tr_vals,
                             int inputs_on_device, mfp_stream s);
int mfp_matrix_view_get(mfp_matrix m, mfp_matrix_view *out);
int mfp_matrix_download(mfp_matrix m, uint32_t *row_ptr, uint32_t *columns, void *values);
int mfp_matrix_destroy(mfp_matrix m);

/* Copy of a matrix with its values converted to dtype (from f32/f64 values;
 * 16-bit types are rounded from the FP32 value, nearest-even).  No reference
 * counterpart: the value-precision study.  Matrices with 16-bit values plan
 * as CSR-vector, ELL, or CSR-Adaptive (rows binned by length, long rows in
 * chunks); all accumulate in FP32 with Dot2 compensation. */
int mfp_matrix_convert(mfp_matrix src, int dtype, mfp_matrix *out);

/* FP64 oracle SpMV on the GPU: y64 = A x with the stored values and x
 * widened to double (one wave per row, double fma), and absdot = sum |a x|
 * per row (may be NULL). */
int mfp_spmv_f64_oracle(mfp_matrix m, int x_dtype, const void *x, double *y64, double *absdot,
                        mfp_stream s);
typedef struct {
  double rel_l2;      /* ||y - y64|| / ||y64|| */
  double max_row_err; /* max |y - y64| / absdot (or / |y64| without absdot) */
  double max_abs_err;
} mfp_error_report;
/* Residual of a device result y (f32 / f64) against device y64 / absdot. */
int mfp_residual(uint32_t n, int y_dtype, const void *y, const double *y64, const double *absdot,
                 mfp_error_report *out);

/* get_rows_statistics (common/matrix_converter.cpp:95-127), on the device. */
typedef struct {
  uint32_t min_elements_in_rows, max_elements_in_rows, avg_elements_in_rows;
  double elements_in_rows_std_deviation;
} mfp_row_stats;
int mfp_matrix_row_stats(mfp_matrix m, mfp_row_stats *out);

/* ---- plans: one format built on the device from a matrix ---------------- */
typedef struct {
  uint32_t lanes;         /* CSR-vector lanes per row (2..64) */
  uint32_t nnz_per_wg;    /* CSR-Adaptive row-block size (csr_adaptive_spmv.cu:8) */
  uint32_t long_row;      /* CSR-Adaptive: rows longer than this are split */
  uint32_t chunk;         /* CSR-Adaptive: elements per split chunk */
  uint32_t ell_width;     /* ELL width (0 = max row length) */
  double hyb_percent;     /* hybrid width rule (matrix_converter.cpp:302), if hyb_width == 0 */
  uint32_t hyb_width;     /* explicit hybrid ELL width (0 = use hyb_percent) */
  uint32_t scoo_slices;   /* SCOO slices_count (0 = MI355X sizing) */
  uint32_t scoo_slice_size;
  uint32_t scoo_lanes;    /* SCOO lane_size */
  uint32_t scoo_threads;  /* SCOO workgroup size */
  uint64_t budget_bytes;  /* refuse layouts above this (0 = 90% of free HBM) */
  uint32_t flags;         /* MFP_TUNE_* */
  uint32_t scoo_pieces;   /* column-sorted SCOO: work items per slice (0 = about two per CU in all) */
} mfp_tuning;
enum {
  MFP_TUNE_NO_OFFSET16 = 1, /* 16-bit-value ELL: keep 32-bit columns */
  MFP_TUNE_IS_DEFAULT = 2,  /* set by the autotuner on each format's default candidate; no effect */
  /* SCOO with each slice's entries sorted by column and split into work
   * items (about scoo_pieces per slice; 0 = about two per CU over the whole
   * matrix), none spanning 2^(32 - log2 scoo_slice_size) columns or more
   * (a wider slice is cut at the multiples of that width); scoo_slice_size
   * rows (a power of two, at most 16384: 8 bytes
   * of LDS per row) accumulate in LDS, and rows much longer than a slice's
   * mean share get slices of their own.  No reference counterpart (the
   * reference's SCOO keeps row order, scoo_matrix_class.cpp:102). */
  MFP_TUNE_SCOO_COLSORT = 4,
  /* The plan gets no buffer of partial results of its own (column-sorted
   * SCOO partial slots, CSR-Adaptive long-row chunk sums, chunk sums of the
   * binned 16-bit CSR): such a plan runs through mfp_plan_spmv_noalloc, and
   * mfp_plan_spmv / mfp_plan_spmm refuse it (MFP_ERR_STATE) when it has
   * partial results. */
  MFP_TUNE_STATIC_PARTIALS = 8,
  /* With MFP_TUNE_SCOO_COLSORT: no long-row slices and one work item per
   * slice (more only where a slice's columns span a band or more), so
   * scoo_pieces is ignored.  A layout whose every slice is one item has no
   * partial slots and runs as one kernel.  Meant for many small matrices
   * in one block-diagonal batch, where slices are many and narrow. */
  MFP_TUNE_SCOO_WHOLE_SLICES = 16,
  /* FP32 and FP64 ELL plans: a layout of 64-row slices with 16-bit column
   * offsets from each slice's first column, four slots to a word (four
   * values, 8 bytes of offsets): 6 bytes per slot instead of 8 (FP64: 10
   * instead of 12), summed in the same slot order as the plain ELL.  Refused
   * (MFP_ERR_ARG) for 16-bit values and when a slice's columns span 65,536
   * or more.  Such a plan keeps no column-major
   * arrays, so mfp_plan_array has none to give for it.  No reference
   * counterpart (the reference's ELL is column-major, matrix_converter.cpp:170). */
  MFP_TUNE_ELL_PACKED = 32
};

typedef struct mfp_plan_s *mfp_plan;
int mfp_tuning_default(mfp_matrix m, int format, mfp_tuning *out);
int mfp_plan_create(mfp_plan *out, mfp_matrix m, int format, const mfp_tuning *t,
                    mfp_stream s);
/* y = A x on stream s (no host sync; graph-capturable). */
int mfp_plan_spmv(mfp_plan p, const void *x, void *y, mfp_stream s);
/* One mfp_plan_spmv with an event recorded before it and after each kernel
 * it launches; waits for the last event.  ms[i] (i < min(*n_kernels,
 * max_kernels)) is kernel i's own duration where its launcher records the
 * kernel's start and end (hipExtLaunchKernelGGL events: every FP32 / FP64
 * format), else the time from the previous event to the end of kernel i
 * (16-bit value plans, and the zero fill of y that COO and the atomic
 * hybrid launch first); *n_kernels is the number of kernels (at most 8; 1
 * with an empty interval when nothing is launched).  A plan's SpMV is one
 * kernel for most formats; two for COO and the atomic hybrid (fill, kernel),
 * the 2-kernel hybrid (ELL, COO), the column-sorted SCOO (main, finish) and
 * CSR-Adaptive with long-row chunks.  Splits the event-pair kernel timing of
 * gpu_matrix_multiplier.cu:92-112 per kernel; not for graph capture, not
 * re-entrant on one host thread. */
int mfp_plan_kernel_times(mfp_plan p, const void *x, void *y, mfp_stream s, float *ms, int max_kernels,
                          int *n_kernels);
/* mfp_plan_spmv that allocates nothing and needs no buffer of the plan's
 * own: the partial results its kernels write and read back (column-sorted
 * SCOO partial slots, CSR-Adaptive long-row chunk sums, chunk sums of the
 * binned 16-bit CSR) go to a static device array of
 * MFP_NOALLOC_SCRATCH_BYTES per device instead (held by each device that
 * runs it until the process ends).  A plan whose partial results need more
 * gets MFP_ERR_BUDGET.  Every partial result is written before it is read
 * within one call, so nothing in the array has to start at zero.  All calls
 * on one device share the array: calls that overlap in time on one device
 * (two streams, two host threads) race.  No counterpart in the reference;
 * for callers that cannot allocate (HIP graph capture, tight serving loops)
 * and for holding many plans created with MFP_TUNE_STATIC_PARTIALS. */
#define MFP_NOALLOC_SCRATCH_BYTES (32u << 20)
int mfp_plan_spmv_noalloc(mfp_plan p, const void *x, void *y, mfp_stream s);
/* Y = A X for k right-hand sides (k in {1,2,4,8,16}); X is cols x k and Y
 * rows x k, row-major; X is 16-byte aligned when k values make a multiple
 * of 16 bytes.  k > 1 runs on CSR-vector plans (f32 / f64; each column is
 * summed in the CSR-vector order), on CSR-Adaptive plans (f32 / f64, the
 * CSR-stream SpMM of MFP_BATCH_STREAM over row blocks of 512 entries; k
 * values must fill whole 16-byte pieces, X and Y 16-byte aligned) and on
 * column-sorted SCOO plans (f32 / f64, while scoo_slice_size * k doubles fit
 * in the 160 KB LDS: k <= 4 for 4096-row slices, 16 for 1024; exact
 * products added in double per slice row and column, partial slots added in
 * slot order).  No reference counterpart (the reference multiplies one
 * vector). */
int mfp_plan_spmm(mfp_plan p, uint32_t k, const void *X, void *Y, mfp_stream s);

/* ---- batches: many matrices in one launch ----------------------------------
 * No reference counterpart.  The matrices (same value type, f32 / f64) are
 * assembled on the device into one block-diagonal CSR; X and Y are the
 * matrices' vectors stacked in order (row i of matrix j at row_offsets[j] + i
 * of Y, column offsets likewise for X), k columns each, row-major. */
typedef struct mfp_batch_s *mfp_batch;
typedef struct mfp_graph_s *mfp_graph;
/* lanes, for k > 1: CSR-vector lanes per row (2..64); 0 to give each row the
 * next power of two of its length, from 2 to 64 lanes (each column is then
 * summed in the CSR-vector order with its row's lane count); 1 for the
 * column-sorted SCOO SpMM over the assembled matrix (one plan per k, built by
 * the first mfp_batch_spmm or mfp_batch_graph_create with that k: the
 * largest power-of-two slice whose k doubles per row fit in 160 KB of LDS,
 * one work item per slice and column band, 1,024 threads); MFP_BATCH_STREAM
 * for the CSR-stream SpMM: row blocks of whole rows with at most 512 entries
 * (fill_row_blocks, a longer row alone), staged in LDS, each group of lanes
 * that covers one X row summing 512 / groups consecutive entries, rows that
 * cross groups finished by the group of their last entry (k values must fill
 * whole 16-byte pieces: k >= 4 for f32, >= 2 for f64). */
#define MFP_BATCH_STREAM 3u
/* k1_tuning: the plan for k = 1 over the assembled matrix: CSR-Adaptive
 * (NULL = its defaults), or the column-sorted SCOO when its flags hold
 * MFP_TUNE_SCOO_COLSORT (two kernels, not bit-exact to a row order). */
int mfp_batch_create(mfp_batch *out, const mfp_matrix *matrices, uint32_t count, uint32_t lanes,
                     const mfp_tuning *k1_tuning, mfp_stream s);
/* offsets arrays get count + 1 entries (may be NULL) */
int mfp_batch_info(mfp_batch b, uint32_t *rows, uint32_t *cols, uint64_t *nnz, uint32_t *row_offsets,
                   uint32_t *col_offsets);
/* the assembled block-diagonal matrix (owned by the batch) */
int mfp_batch_matrix(mfp_batch b, mfp_matrix *out);
/* one launch: CSR-Adaptive (or the column-sorted SCOO plan) for k = 1; for
 * k > 1 the SpMM the batch's lanes select */
int mfp_batch_spmm(mfp_batch b, uint32_t k, const void *X, void *Y, mfp_stream s);
/* capture that launch sequence once in a HIP graph; replay with mfp_graph_launch */
int mfp_batch_graph_create(mfp_batch b, uint32_t k, const void *X, void *Y, mfp_graph *out);
/* capture one mfp_plan_spmv(p, x, y) in a HIP graph (the plan, x and y must
 * outlive it); replay with mfp_graph_launch.  No reference counterpart: the
 * reference launches each kernel from the host every call. */
int mfp_plan_graph_create(mfp_plan p, const void *x, void *y, mfp_graph *out);
int mfp_graph_launch(mfp_graph g, mfp_stream s);
int mfp_graph_destroy(mfp_graph g);
int mfp_batch_destroy(mfp_batch b);

typedef struct {
  int format, dtype;
  uint32_t rows, cols;
  uint64_t nnz;
  uint64_t device_bytes;     /* format arrays owned by the plan */
  uint64_t stored_elements;  /* ELL slots / COO entries / ... */
  uint32_t ell_width;
  uint64_t coo_elements;
  uint32_t blocks, chunks;   /* CSR-Adaptive row blocks / long-row chunks */
  uint32_t slices_count, slice_size, lane_size;
  uint32_t lanes;
  uint32_t nnz_per_wg;
  uint32_t column_bytes;     /* bytes per stored column index (2 with 16-bit offsets) */
  uint32_t scoo_colsort;     /* column-sorted SCOO: slices_count counts the long-row slices too,
                                lane_size holds the work items */
  uint32_t scoo_packed;      /* column-sorted SCOO: 1 (each entry packs its column offset and row in one u32) */
  uint32_t scoo_long_rows;   /* column-sorted SCOO: rows moved to long-row slices */
} mfp_plan_info;
int mfp_plan_get_info(mfp_plan p, mfp_plan_info *out);

/* Device arrays of a plan, for layout parity checks. */
enum {
  MFP_ARR_ELL_VALUES = 0, MFP_ARR_ELL_COLS = 1,
  MFP_ARR_COO_VALUES = 2, MFP_ARR_COO_COLS = 3, MFP_ARR_COO_ROWS = 4,
  MFP_ARR_SCOO_SLICES_PTR = 5, MFP_ARR_ROW_BLOCKS = 6
};
int mfp_plan_array(mfp_plan p, int which, const void **dev_ptr, uint64_t *count);
int mfp_plan_destroy(mfp_plan p);

/* ---- autotuner (no reference counterpart) -----------------------------------
 * Candidates sweep lanes per row (CSR-vector), the CSR-Adaptive block size,
 * the ELL/COO split of the hybrids (row-length quantiles) and the SCOO slice
 * size, plus each format's default tuning (marked MFP_TUNE_IS_DEFAULT; the
 * hybrids' default takes its width from hyb_percent).  mfp_autotune_run
 * builds and times each (median of reps event-timed SpMVs, a cache flush
 * before each when flush != 0); a candidate that cannot be built keeps
 * ms = +inf and its status. */
typedef struct {
  int format;
  mfp_tuning tuning;
  double ms;
  int status;
} mfp_candidate;
/* format_mask bit f enables format f; out may be NULL to count */
int mfp_autotune_candidates(mfp_matrix m, uint32_t format_mask, mfp_candidate *out, uint32_t cap,
                            uint32_t *n);
int mfp_autotune_run(mfp_matrix m, const void *x, void *y, mfp_candidate *cands, uint32_t n,
                     uint32_t reps, int flush);

/* ---- learned format selector (no reference counterpart) -------------------
 * mfp_matrix_features writes 64 structure features of a matrix (row-length
 * histograms and moments, distance from the diagonal, row spans; computed on
 * the device from integer counters, so they do not depend on atomic order).
 * The selector network is an MLP 64 -> 128 -> 128 -> 8 (ReLU, softmax over
 * the eight formats, 25,864 FP32 parameters) whose forward, backward
 * (softmax cross-entropy against target distributions) and Adam update run
 * in this library's HIP kernels.  Parameters are packed
 * W1[64][128] b1[128] W2[128][128] b2[128] W3[128][8] b3[8] (W[in][out]).
 * Inference with bf16 != 0 rounds activations and weights to BF16 and
 * multiplies on the matrix cores with FP32 accumulation.  X is n x 64 and
 * P, probs are n x 8, row-major, on the host; loss is the batch mean. */
int mfp_matrix_features(mfp_matrix m, float *out /* 64 */);
typedef struct mfp_mlp_s *mfp_mlp;
int mfp_mlp_create(mfp_mlp *out, uint64_t seed);
int mfp_mlp_destroy(mfp_mlp m);
uint32_t mfp_mlp_param_count(void);
int mfp_mlp_get_params(mfp_mlp m, float *out);
int mfp_mlp_set_params(mfp_mlp m, const float *in); /* also resets Adam state */
int mfp_mlp_loss_grad(mfp_mlp m, const float *X, const float *P, uint32_t n, double *loss,
                      float *grad);
int mfp_mlp_train_step(mfp_mlp m, const float *X, const float *P, uint32_t n, float lr,
                       double *loss);
int mfp_mlp_predict(mfp_mlp m, const float *X, uint32_t n, int bf16, float *probs);

/* ---- host converters run on the device (compat layer) --------------------- */
/* fill_row_blocks (gpu/csr_adaptive_spmv.cu:180-240), nnz_per_wg generalising
 * NNZ_PER_WG; row_ptr on the host; out gets count+1 entries (out may be NULL
 * to query the count). */
int mfp_fill_row_blocks(uint32_t rows, const uint32_t *row_ptr, uint32_t nnz_per_wg,
                        uint32_t *out, uint32_t *count);
/* scoo_matrix_class geometry (common/scoo_matrix_class.cpp:30-62). */
int mfp_scoo_geometry(int dtype, uint32_t rows, uint32_t sm_count, uint64_t shared_mem_size,
                      uint32_t *slices_count, uint32_t *slice_size, uint32_t *lane_size);

/* ---- the reference's entry points --------------------------------------
 * Each call does what the reference's does: grow the caller's buffers, copy
 * the host format in, set x = 1 (and y), time the SpMV with an event pair,
 * copy y into reusable_vector, run compare_results against reference_y
 * (prints to stderr only; NULL skips it) and return the record.  They use
 * this library's kernels. */
typedef struct {
  char format[64];
  double elapsed;                  /* seconds */
  double load_store_bytes;         /* the reference's byte model */
  double operations_count;
  double effective_bandwidth;      /* GB/s, measurement_class.cpp:13 */
  double computational_throughput; /* GFLOP/s, measurement_class.cpp:14 */
} mfp_measurement;

typedef struct {
  uint32_t rows, cols;
  uint64_t nnz;
  const uint32_t *row_ptr, *columns;
  const void *data;
} mfp_host_csr;

typedef struct {
  uint32_t rows, cols;
  uint64_t nnz;
  uint32_t elements_in_rows;
  const uint32_t *columns;
  const void *data;
} mfp_host_ell;

typedef struct {
  uint32_t rows, cols;
  uint64_t nnz;      /* matrix nnz (meta.non_zero_count) */
  uint64_t elements; /* stored entries (get_matrix_size) */
  const uint32_t *cols_, *rows_;
  const void *data;
} mfp_host_coo;

typedef struct {
  uint32_t rows, cols;
  uint64_t nnz;
  uint32_t slices_count, slice_size, lane_size;
  const void *values;
  const uint32_t *c_index, *r_index, *slices_ptr;
} mfp_host_scoo;

/* compare_results<T> (common/measurement_class.h:60-89): prints the relative
 * squared error and the first differing entry to stderr when the error
 * exceeds 1e-9; never fails.  dtype MFP_F32 or MFP_F64. */
int mfp_compare_results(int dtype, uint32_t n, const void *a, const void *b);

/* gpu_csr_spmv (gpu/gpu_matrix_multiplier.h:15-25, .cu:54-133) */
int mfp_gpu_csr_spmv(int dtype, const mfp_host_csr *A, mfp_buffer A_dev, mfp_buffer col_ids,
                     mfp_buffer row_ptr, mfp_buffer x, mfp_buffer y, void *reusable_vector,
                     const void *reference_y, mfp_measurement *out);
/* gpu_csr_vector_spmv (gpu/gpu_matrix_multiplier.h:39-49, .cu:364-447) */
int mfp_gpu_csr_vector_spmv(int dtype, const mfp_host_csr *A, mfp_buffer A_dev,
                            mfp_buffer col_ids, mfp_buffer row_ptr, mfp_buffer x,
                            mfp_buffer y, void *reusable_vector, const void *reference_y,
                            mfp_measurement *out);
/* gpu_csr_adaptive_spmv (gpu/csr_adaptive_spmv.h:13-23, .cu:242-321) */
int mfp_gpu_csr_adaptive_spmv(int dtype, const mfp_host_csr *A, mfp_buffer A_dev,
                              mfp_buffer col_ids, mfp_buffer row_ptr, mfp_buffer x,
                              mfp_buffer y, void *reusable_vector, const void *reference_y,
                              mfp_measurement *out);
/* gpu_ell_spmv (gpu/gpu_matrix_multiplier.h:64-73, .cu:473-550) */
int mfp_gpu_ell_spmv(int dtype, const mfp_host_ell *A, mfp_buffer A_dev, mfp_buffer col_ids,
                     mfp_buffer x, mfp_buffer y, void *reusable_vector,
                     const void *reference_y, mfp_measurement *out);
/* gpu_coo_spmv (gpu/gpu_matrix_multiplier.h:87-97, .cu:640-716) */
int mfp_gpu_coo_spmv(int dtype, const mfp_host_coo *A, mfp_buffer A_dev, mfp_buffer col_ids,
                     mfp_buffer row_ids, mfp_buffer x, mfp_buffer y, void *reusable_vector,
                     const void *reference_y, mfp_measurement *out);
/* gpu_scoo_spmv (gpu/scoo_spmv.h:12-23, .cu:144-222) */
int mfp_gpu_scoo_spmv(int dtype, int print_diff, const mfp_host_scoo *A, mfp_buffer A_dev,
                      mfp_buffer r_index, mfp_buffer c_index, mfp_buffer x, mfp_buffer y,
                      void *reusable_vector, const void *reference_y, mfp_measurement *out);
/* gpu_hybrid_spmv (gpu/gpu_matrix_multiplier.h:112-124, .cu:866-967) */
int mfp_gpu_hybrid_spmv(int dtype, const mfp_host_ell *ell, const mfp_host_coo *coo,
                        mfp_buffer A_ell, mfp_buffer A_coo, mfp_buffer ell_col_ids,
                        mfp_buffer coo_col_ids, mfp_buffer coo_row_ids, mfp_buffer x,
                        mfp_buffer y, void *reusable_vector, const void *reference_y,
                        mfp_measurement *out);
/* gpu_hybrid_atomic_spmv (gpu/gpu_matrix_multiplier.h:126-138, .cu:1004-1096) */
int mfp_gpu_hybrid_atomic_spmv(int dtype, const mfp_host_ell *ell, const mfp_host_coo *coo,
                               mfp_buffer A_ell, mfp_buffer A_coo, mfp_buffer ell_col_ids,
                               mfp_buffer coo_col_ids, mfp_buffer coo_row_ids, mfp_buffer x,
                               mfp_buffer y, void *reusable_vector, const void *reference_y,
                               mfp_measurement *out);
/* gpu_hybrid_cpu_coo_spmv (gpu/gpu_matrix_multiplier.h:141-152, .cu:1111-1217):
 * the ELL part on the GPU while the host sums the COO tail into cpu_y (rows
 * entries, caller-owned, zeroed here); the tail is copied into tmp and added
 * to y.  The timed span runs from the ELL launch to the end of the add, host
 * loop included.  The reference's driver keeps this row switched off
 * (main.cpp:247); mfp_bench runs it only when asked. */
int mfp_gpu_hybrid_cpu_coo_spmv(int dtype, const mfp_host_ell *ell, const mfp_host_coo *coo,
                                mfp_buffer A_ell, mfp_buffer ell_col_ids, mfp_buffer x,
                                mfp_buffer y, mfp_buffer tmp, void *cpu_y, void *reusable_vector,
                                const void *reference_y, mfp_measurement *out);

/* ---- synthetic matrices (host generator, seeded, deterministic) ----------- */
enum {
  MFP_GEN_STENCIL2D = 0, /* 5-point, n x n grid */
  MFP_GEN_STENCIL3D = 1, /* 27-point, n x n x n grid */
  MFP_GEN_POWERLAW = 2,  /* nnz/row = min(max, floor(1/U)) (Pareto density x^-2) */
  MFP_GEN_UNIFORM = 3    /* nnz/row uniform in [lo, hi] */
};
enum { MFP_VAL_LAPLACE = 0, MFP_VAL_UNIFORM = 1, MFP_VAL_NORMAL = 2 };
typedef struct {
  int family;
  int values;            /* MFP_VAL_* */
  uint64_t seed;
  uint32_t n;            /* stencil grid edge */
  uint32_t rows;         /* powerlaw / uniform */
  uint32_t nnz_lo, nnz_hi; /* powerlaw min/max row length, uniform range */
  double alpha;          /* powerlaw density exponent (2.0 uses the exact 1/U form) */
  double bandwidth;      /* column window as a fraction of cols (1 = full) */
} mfp_gen_params;
/* Pass NULL arrays to get rows/nnz first. */
int mfp_generate(const mfp_gen_params *p, uint32_t *rows, uint32_t *cols, uint64_t *nnz,
                 uint32_t *row_ptr, uint32_t *columns, void *values, int dtype);
/* x ~ U[-1,1) from seed (host). */
int mfp_generate_vector(uint64_t seed, uint64_t n, void *x, int dtype);

#ifdef __cplusplus
}
#endif

#endif
