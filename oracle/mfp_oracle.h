/*
 * CPU oracle for the SpMV hot path.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's host converters, CPU SpMV
 * functions, partition builders and measurement arithmetic.  It exists to
 * check the HIP library; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and never as the thing measured or shipped.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the reference repository root).  The pins that check each function are
 * listed in oracle/README.md.
 */
#ifndef MFP_ORACLE_H
#define MFP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_F32 0
#define ORC_F64 1

/* ---- L0 converters (common/matrix_converter.cpp) ---------------------- */

/* Triplets -> CSR.  common/matrix_converter.cpp:15-87: count per row
 * (:35-36), exclusive scan (:38-44), scatter in input order (:46-57), per-row
 * sort by column (:59-85).  Ties (duplicate (row,col)) keep input order
 * (stable); the reference's std::sort leaves that order unspecified.
 * Returns 0, or -1 when a triplet row/col is out of range. */
int orc_csr_from_triplets(uint32_t rows, uint32_t cols, uint64_t nnz,
                          const uint32_t *tr_rows, const uint32_t *tr_cols,
                          const double *tr_vals, int dtype,
                          uint32_t *row_ptr, uint32_t *columns, void *data);

/* common/matrix_converter.cpp:95-127.  out3 = {min, max, avg} where avg is
 * the 32-bit integer quotient sum/rows (:116); std uses that integer avg. */
void orc_rows_statistics(uint32_t rows, const uint32_t *row_ptr,
                         uint32_t out3[3], double *out_std);

/* ELL with a given width; rows longer than width are truncated.
 * common/matrix_converter.cpp:176-203 (width = max row length gives the
 * :139-174 constructor).  Column-major data[row + k*rows], pad (0, col 0). */
void orc_ell_from_csr(uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data, int dtype,
                      uint32_t width, void *ell_data, uint32_t *ell_cols);

/* Number of COO entries whose in-row offset is >= element_start
 * (common/matrix_converter.cpp:242-255). */
uint64_t orc_coo_count(uint32_t rows, const uint32_t *row_ptr,
                       uint32_t element_start);

/* COO in CSR order, entries with in-row offset >= element_start.
 * element_start = 0 is the :211-240 constructor, >0 the :242-280 one. */
void orc_coo_from_csr(uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data, int dtype,
                      uint32_t element_start, void *coo_data,
                      uint32_t *coo_cols, uint32_t *coo_rows);

/* Hybrid ELL width (common/matrix_converter.cpp:296-302):
 * (unsigned)(avg + (max - avg) * percent). */
uint32_t orc_hybrid_width(uint32_t rows, const uint32_t *row_ptr,
                          double percent);

/* ---- SCOO (common/scoo_matrix_class.cpp) -------------------------------- */

uint32_t orc_next_pow_2(uint32_t x); /* :13-21 */
uint32_t orc_prev_pow_2(uint32_t x); /* :23-28 */
/* :30-44 */
uint32_t orc_scoo_slices_count(uint32_t rows, uint32_t sm_count,
                               uint64_t shared_mem_size, int dtype);
/* :46-56 */
uint32_t orc_scoo_lane_size(uint32_t slices_count, uint32_t slice_size,
                            uint64_t shared_mem_size, int dtype);
/* :58-101.  values/c_index/r_index in CSR order; slices_ptr has
 * slices_count+1 entries, the ones after the last real slice stay 0. */
void orc_scoo_build(uint32_t rows, const uint32_t *row_ptr,
                    const uint32_t *columns, const void *data, int dtype,
                    uint32_t slices_count, uint32_t slice_size,
                    void *values, uint32_t *c_index, uint32_t *r_index,
                    uint32_t *slices_ptr);

/* ---- CSR-Adaptive row blocks (gpu/csr_adaptive_spmv.cu:180-240) -------- */

/* nnz_per_wg generalises the reference's NNZ_PER_WG (64, :8).  With
 * fill == 0 only counts.  Returns the block count; row_blocks gets
 * count + 1 entries. */
uint32_t orc_fill_row_blocks(int fill, uint32_t rows, const uint32_t *row_ptr,
                             uint32_t *row_blocks, uint32_t nnz_per_wg);

/* ---- SpMV --------------------------------------------------------------- */

/* cpu/cpu_matrix_multiplier.cpp:20-61: sequential dot per row in the
 * value type.  The reference is built -O3 -march=native (cpu/CMakeLists.txt:7)
 * in GNU mode, where GCC contracts `dot += a*x` into a fused multiply-add; the
 * restatement therefore calls fma()/fmaf() explicitly. */
void orc_spmv_csr_seq(int dtype, uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data,
                      const void *x, void *y);

/* Same loop with a separate rounded multiply and add (nvcc -fmad=false /
 * debug builds, gpu/CMakeLists.txt:17-19). */
void orc_spmv_csr_seq_nofma(int dtype, uint32_t rows, const uint32_t *row_ptr,
                            const uint32_t *columns, const void *data,
                            const void *x, void *y);

/* Reduction-order emulator.  cpu/cpu_matrix_multiplier.cpp:170-227 with the
 * 32 lanes generalised to `lanes` (a power of two <= 64): lane l accumulates
 * elements l, l+lanes, ... with fma, then a halving tree
 * v[i] += v[i + off] for off = lanes/2 .. 1.  lanes == 32 is the reference. */
void orc_spmv_csr_lanes(int dtype, uint32_t lanes, uint32_t rows,
                        const uint32_t *row_ptr, const uint32_t *columns,
                        const void *data, const void *x, void *y);

/* Host COO loop of the host-COO hybrid (gpu/gpu_matrix_multiplier.cu:
 * 1163-1168): y[row[e]] += data[e] * x[col[e]] in stored order, the product
 * rounded before the add; y is accumulated into. */
void orc_spmv_coo_seq(int dtype, uint64_t n, const uint32_t *coo_rows,
                      const uint32_t *coo_cols, const void *coo_data,
                      const void *x, void *y);

/* Summation order of this library's CSR-stream SpMM, one column of X
 * (batch.hip k_spmm_stream): row blocks of fill_row_blocks(W); a block with
 * more than W entries is one row summed by `groups` lane groups (group g:
 * entries g, g + groups, ... with fma) and a halving tree; otherwise a row's
 * segment in each group of W / groups consecutive entries is an fma sum, and
 * a row across groups ga..gb is seg[gb] + seg[gb-1] + ... + seg[ga].  Returns
 * 0, or -1 for a bad W / groups. */
int orc_spmm_stream_col(int dtype, uint32_t W, uint32_t groups, uint32_t rows,
                        const uint32_t *row_ptr, const uint32_t *columns,
                        const void *data, const void *x, void *y);

/* ELL: thread per row, sequential over the padded width with fma
 * (gpu/gpu_matrix_multiplier.cu:450-471). */
void orc_spmv_ell_seq(int dtype, uint32_t rows, uint32_t width,
                      const uint32_t *ell_cols, const void *ell_data,
                      const void *x, void *y);

/* Summation order of this library's CSR-Adaptive kernel (not the
 * reference's): blocks from orc_fill_row_blocks(nnz_per_wg); multi-row
 * blocks stage rounded products a*x, then each row is summed by t threads
 * strided by t (t = largest power of two <= min(64, block_threads/rows))
 * followed by a halving tree; t == 1 sums sequentially.  Single-row blocks:
 * block_threads lanes strided with fma, 64-lane halving tree per wave, then a
 * halving tree over the waves.  Rows longer than long_row are split into
 * chunks of `chunk` elements, each reduced as a single-row block, and the
 * chunk partials are added left to right. */
void orc_spmv_csr_adaptive(int dtype, uint32_t rows, const uint32_t *row_ptr,
                           const uint32_t *columns, const void *data,
                           const void *x, void *y, uint32_t nnz_per_wg,
                           uint32_t block_threads, uint32_t long_row,
                           uint32_t chunk);

/* FP64 truth: y64 = A*x evaluated in double (values and x widened from
 * dtype), plus absdot[r] = sum_j |a_rj * x_j| for per-row normalisation. */
void orc_spmv_csr_f64(int dtype, uint32_t rows, const uint32_t *row_ptr,
                      const uint32_t *columns, const void *data,
                      const void *x, double *y64, double *absdot);

/* cpu/cpu_matrix_multiplier.cpp:229-308: static row split over `threads`
 * threads (the last one takes the remainder), spin-barrier start, each
 * thread times its own rows; returns the maximum thread time in seconds. */
double orc_spmv_csr_multi_thread(int dtype, uint32_t threads, uint32_t rows,
                                 const uint32_t *row_ptr,
                                 const uint32_t *columns, const void *data,
                                 const void *x, void *y);

/* Wall time of `reps` single-thread sequential passes (seconds per pass). */
double orc_spmv_csr_seq_timed(int dtype, uint32_t reps, uint32_t rows,
                              const uint32_t *row_ptr, const uint32_t *columns,
                              const void *data, const void *x, void *y);

/* ---- measurement (common/measurement_class.{h,cpp}) --------------------- */

/* measurement_class.h:60-72: sum((a-b)^2) / sum(b^2) accumulated in the
 * value type. */
double orc_compare_error(int dtype, uint32_t n, const void *a, const void *b);

/* measurement_class.cpp:7-16: out2 = {bytes/(elapsed*1e9), ops/(elapsed*1e9)} */
void orc_measurement(double elapsed, double bytes, double ops, double out2[2]);

/* Reference byte models (A15).  kind: 0 CSR family
 * (gpu_matrix_multiplier.cu:120-124), 1 ELL (:537-541, n = width*rows taken
 * modulo 2^32 as the reference's unsigned int does), 2 COO (:704-708). */
double orc_reference_bytes(int kind, int dtype, uint32_t rows, uint64_t nnz,
                           uint32_t ell_width);

#ifdef __cplusplus
}
#endif

#endif
