// Benchmark-only rocSPARSE CSR SpMV slot (see rocsparse_slot.h).  The
// rocSPARSE types are mirrored here as plain integers and opaque pointers so
// that no rocSPARSE header or library is needed to build it; the values are
// those of rocsparse-types.h (ROCm 7.2).
#include "rocsparse_slot.h"

#include "../../include/mfp.h"

#include <dlfcn.h>

#include <mutex>
#include <string>
#include <type_traits>

namespace {

using rs_status = int;
using rs_handle = void *;
using rs_spmat = void *;
using rs_dnvec = void *;
constexpr int RS_SUCCESS = 0;
constexpr int RS_OP_NONE = 111;        // rocsparse_operation_none
constexpr int RS_BASE_ZERO = 0;        // rocsparse_index_base_zero
constexpr int RS_I32 = 2;              // rocsparse_indextype_i32
constexpr int RS_F32 = 151;            // rocsparse_datatype_f32_r
constexpr int RS_F64 = 152;            // rocsparse_datatype_f64_r
constexpr int RS_ALG_CSR_ADAPTIVE = 2; // rocsparse_spmv_alg_csr_adaptive
constexpr int RS_STAGE_BUFFER_SIZE = 1, RS_STAGE_PREPROCESS = 2, RS_STAGE_COMPUTE = 3;

struct api {
  void *lib = nullptr;
  rs_status (*create_handle)(rs_handle *) = nullptr;
  rs_status (*destroy_handle)(rs_handle) = nullptr;
  rs_status (*set_stream)(rs_handle, void *) = nullptr;
  rs_status (*create_csr_descr)(rs_spmat *, int64_t, int64_t, int64_t, void *, void *, void *, int, int, int,
                                int) = nullptr;
  rs_status (*destroy_spmat_descr)(rs_spmat) = nullptr;
  rs_status (*create_dnvec_descr)(rs_dnvec *, int64_t, void *, int) = nullptr;
  rs_status (*destroy_dnvec_descr)(rs_dnvec) = nullptr;
  rs_status (*spmv)(rs_handle, int, const void *, rs_spmat, rs_dnvec, const void *, rs_dnvec, int, int, int,
                    size_t *, void *) = nullptr;
  std::string error;
  bool ok = false;
};

api &the_api()
{
  static api a;
  static std::once_flag once;
  std::call_once(once, [] {
    a.lib = dlopen("librocsparse.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!a.lib) a.lib = dlopen("librocsparse.so", RTLD_NOW | RTLD_LOCAL);
    if (!a.lib) {
      const char *e = dlerror();
      a.error = std::string("librocsparse.so not found: ") + (e ? e : "");
      return;
    }
    auto sym = [&](auto &fn, const char *name) {
      fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(a.lib, name));
      if (!fn && a.error.empty()) a.error = std::string("librocsparse.so lacks ") + name;
    };
    sym(a.create_handle, "rocsparse_create_handle");
    sym(a.destroy_handle, "rocsparse_destroy_handle");
    sym(a.set_stream, "rocsparse_set_stream");
    sym(a.create_csr_descr, "rocsparse_create_csr_descr");
    sym(a.destroy_spmat_descr, "rocsparse_destroy_spmat_descr");
    sym(a.create_dnvec_descr, "rocsparse_create_dnvec_descr");
    sym(a.destroy_dnvec_descr, "rocsparse_destroy_dnvec_descr");
    sym(a.spmv, "rocsparse_spmv");
    a.ok = a.error.empty();
  });
  return a;
}

thread_local std::string g_err;

} // namespace

struct mfp_rs_plan_s {
  rs_handle h = nullptr;
  rs_spmat A = nullptr;
  rs_dnvec x = nullptr, y = nullptr;
  void *buffer = nullptr;
  size_t buffer_size = 0;
  int dtype = MFP_F32;
};

extern "C" {

int mfp_rs_available(void) { return the_api().ok ? 1 : 0; }

const char *mfp_rs_error(void)
{
  if (!the_api().ok) return the_api().error.c_str();
  return g_err.c_str();
}

int mfp_rs_plan_destroy(mfp_rs_plan p)
{
  if (!p) return MFP_OK;
  api &a = the_api();
  if (a.ok) {
    if (p->x) a.destroy_dnvec_descr(p->x);
    if (p->y) a.destroy_dnvec_descr(p->y);
    if (p->A) a.destroy_spmat_descr(p->A);
    if (p->h) a.destroy_handle(p->h);
  }
  if (p->buffer) mfp_free(p->buffer);
  delete p;
  return MFP_OK;
}

int mfp_rs_plan_create(mfp_rs_plan *out, int dtype, uint32_t rows, uint32_t cols, uint64_t nnz,
                       const uint32_t *row_ptr, const uint32_t *columns, const void *values, const void *x,
                       void *y, void *stream)
{
  api &a = the_api();
  if (!a.ok) {
    g_err = a.error;
    return MFP_ERR_STATE;
  }
  if (!out || (dtype != MFP_F32 && dtype != MFP_F64) || nnz >= (1ull << 31)) {
    g_err = "rocSPARSE slot: f32 / f64 values and fewer than 2^31 entries";
    return MFP_ERR_ARG;
  }
  auto *p = new mfp_rs_plan_s;
  p->dtype = dtype;
  const int dt = dtype == MFP_F32 ? RS_F32 : RS_F64;
  auto fail = [&](const char *what, int code) {
    g_err = std::string("rocSPARSE slot: ") + what;
    mfp_rs_plan_destroy(p);
    return code;
  };
  if (a.create_handle(&p->h) != RS_SUCCESS) return fail("rocsparse_create_handle", MFP_ERR_HIP);
  if (a.set_stream(p->h, stream) != RS_SUCCESS) return fail("rocsparse_set_stream", MFP_ERR_HIP);
  if (a.create_csr_descr(&p->A, rows, cols, (int64_t)nnz, const_cast<uint32_t *>(row_ptr),
                         const_cast<uint32_t *>(columns), const_cast<void *>(values), RS_I32, RS_I32, RS_BASE_ZERO,
                         dt) != RS_SUCCESS)
    return fail("rocsparse_create_csr_descr", MFP_ERR_HIP);
  if (a.create_dnvec_descr(&p->x, cols, const_cast<void *>(x), dt) != RS_SUCCESS ||
      a.create_dnvec_descr(&p->y, rows, y, dt) != RS_SUCCESS)
    return fail("rocsparse_create_dnvec_descr", MFP_ERR_HIP);
  const double one = 1.0, zero = 0.0;
  const float onef = 1.0f, zerof = 0.0f;
  const void *alpha = dtype == MFP_F32 ? (const void *)&onef : (const void *)&one;
  const void *beta = dtype == MFP_F32 ? (const void *)&zerof : (const void *)&zero;
  if (a.spmv(p->h, RS_OP_NONE, alpha, p->A, p->x, beta, p->y, dt, RS_ALG_CSR_ADAPTIVE, RS_STAGE_BUFFER_SIZE,
             &p->buffer_size, nullptr) != RS_SUCCESS)
    return fail("rocsparse_spmv (buffer size)", MFP_ERR_HIP);
  if (p->buffer_size && mfp_malloc(&p->buffer, p->buffer_size) != MFP_OK)
    return fail("buffer allocation", MFP_ERR_HIP);
  if (a.spmv(p->h, RS_OP_NONE, alpha, p->A, p->x, beta, p->y, dt, RS_ALG_CSR_ADAPTIVE, RS_STAGE_PREPROCESS,
             &p->buffer_size, p->buffer) != RS_SUCCESS)
    return fail("rocsparse_spmv (preprocess)", MFP_ERR_HIP);
  *out = p;
  return MFP_OK;
}

int mfp_rs_plan_spmv(mfp_rs_plan p, void *stream)
{
  api &a = the_api();
  if (!p || !a.ok) return MFP_ERR_STATE;
  if (a.set_stream(p->h, stream) != RS_SUCCESS) return MFP_ERR_HIP;
  const double one = 1.0, zero = 0.0;
  const float onef = 1.0f, zerof = 0.0f;
  const bool f32 = p->dtype == MFP_F32;
  const int dt = f32 ? RS_F32 : RS_F64;
  size_t bs = p->buffer_size;
  if (a.spmv(p->h, RS_OP_NONE, f32 ? (const void *)&onef : (const void *)&one, p->A, p->x,
             f32 ? (const void *)&zerof : (const void *)&zero, p->y, dt, RS_ALG_CSR_ADAPTIVE, RS_STAGE_COMPUTE, &bs,
             p->buffer) != RS_SUCCESS) {
    g_err = "rocSPARSE slot: rocsparse_spmv (compute)";
    return MFP_ERR_HIP;
  }
  return MFP_OK;
}

} // extern "C"
