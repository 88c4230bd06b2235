/*
 * Benchmark-only comparator: CSR SpMV by rocSPARSE, the counterpart of the
 * reference's "GPU CSR (cuSparse)" row (gpu/gpu_matrix_multiplier.cu:175-257,
 * main.cpp:194-197).  librocsparse.so is opened at run time (dlopen), so it
 * is never a link dependency of libmfp or of the driver, and no SpMV of the
 * library goes through it.  When the library is absent the slot reports
 * itself unavailable and its row is left out.
 */
#ifndef MFP_ROCSPARSE_SLOT_H
#define MFP_ROCSPARSE_SLOT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 1 when librocsparse.so loads and exports the entry points the slot uses,
 * else 0 (mfp_rs_error says why). */
int mfp_rs_available(void);
const char *mfp_rs_error(void);

/* y = A x for one device-resident CSR matrix (u32 row_ptr / columns read as
 * rocSPARSE's int32 indices; dtype MFP_F32 or MFP_F64) with rocsparse_spmv's
 * CSR-adaptive algorithm.  Create runs its buffer-size and preprocess
 * stages (blocking); spmv enqueues the compute stage on the stream. */
typedef struct mfp_rs_plan_s *mfp_rs_plan;
int mfp_rs_plan_create(mfp_rs_plan *out, int dtype, uint32_t rows, uint32_t cols, uint64_t nnz,
                       const uint32_t *row_ptr, const uint32_t *columns, const void *values, const void *x,
                       void *y, void *stream);
int mfp_rs_plan_spmv(mfp_rs_plan p, void *stream);
int mfp_rs_plan_destroy(mfp_rs_plan p);

#ifdef __cplusplus
}
#endif

#endif
