// Seeded synthetic matrices for the benchmark workloads.
//
// Every random draw is a counter-based hash of (seed, stream, index), so a
// row's length, columns and values do not depend on how rows are split over
// threads.  Families:
//   stencil2d  5-point on an n x n grid (columns ascending: r-n, r-1, r, r+1, r+n)
//   stencil3d  27-point on an n^3 grid (neighbours in (dk, di, dj) order)
//   powerlaw   nnz/row = min(hi, max(lo, floor(lo / U^(1/(alpha-1))))); for
//              alpha = 2 this is floor(lo / U) evaluated exactly
//   uniform    nnz/row uniform in [lo, hi]
// Columns of powerlaw/uniform rows: jittered-stratified without duplicates --
// the row's window of B columns is cut into len equal strata and one column
// is drawn uniformly in each, which keeps them sorted and distinct.
// Values: Laplacian constants, U[-1,1) on a 2^-23 grid, or N(0,1) approximated
// by the Irwin-Hall sum of 12 uniforms (exact in double).
#include "../../include/mfp.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t idx)
{
  return mix64(mix64(seed + 0x632BE59BD9B4E019ull * (stream + 1)) ^ idx);
}

enum : uint64_t { S_LEN = 0, S_COL = 1, S_VAL = 2, S_VEC = 3 };

// U[-1, 1) on a 2^-23 grid: exact in float.
inline double u_pm1(uint64_t h) { return ((double)((h >> 40) & 0xFFFFFF) - 8388608.0) / 8388608.0; }
// (0, 1] with 53 bits.
inline double u_01(uint64_t h) { return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0); }

double draw_value(int mode, uint64_t seed, uint64_t e, bool diag, double diag_val)
{
  switch (mode) {
  case MFP_VAL_LAPLACE: return diag ? diag_val : -1.0;
  case MFP_VAL_UNIFORM: return u_pm1(hash3(seed, S_VAL, e));
  default: {
    double s = 0.0;
    for (uint64_t j = 0; j < 12; j++) s += (double)((hash3(seed, S_VAL, e * 12 + j) >> 40) & 0xFFFFFF) / 16777216.0;
    return s - 6.0;
  }
  }
}

void store_value(void *values, int dtype, uint64_t e, double v)
{
  switch (dtype) {
  case MFP_F32: ((float *)values)[e] = (float)v; break;
  case MFP_F64: ((double *)values)[e] = v; break;
  case MFP_BF16: {
    __hip_bfloat16 b = __float2bfloat16((float)v);
    std::memcpy((uint16_t *)values + e, &b, 2);
    break;
  }
  case MFP_F16: {
    __half h = __float2half((float)v);
    std::memcpy((uint16_t *)values + e, &h, 2);
    break;
  }
  }
}

uint32_t row_len(const mfp_gen_params &p, uint32_t cols, uint32_t r)
{
  const uint64_t h = hash3(p.seed, S_LEN, r);
  uint64_t len;
  const uint32_t lo = std::max<uint32_t>(1, p.nnz_lo);
  const uint32_t hi = std::max<uint32_t>(lo, p.nnz_hi);
  if (p.family == MFP_GEN_POWERLAW) {
    const double u = u_01(h);
    double x;
    if (p.alpha == 2.0)
      x = (double)lo / u;
    else
      x = (double)lo * std::pow(u, -1.0 / (p.alpha - 1.0));
    len = x >= (double)hi ? hi : (uint64_t)std::floor(x);
    len = std::max<uint64_t>(len, lo);
  } else {
    len = lo + h % ((uint64_t)hi - lo + 1);
  }
  return (uint32_t)std::min<uint64_t>(len, cols);
}

void window(const mfp_gen_params &p, uint32_t rows, uint32_t cols, uint32_t r, uint32_t len,
            uint64_t *w0, uint64_t *B)
{
  uint64_t b = (p.bandwidth >= 1.0 || p.bandwidth <= 0.0) ? cols : (uint64_t)std::llround(p.bandwidth * cols);
  b = std::min<uint64_t>(cols, std::max<uint64_t>(b, len));
  const uint64_t center = rows ? (uint64_t)r * cols / rows : 0;
  int64_t start = (int64_t)center - (int64_t)(b / 2);
  start = std::max<int64_t>(0, std::min<int64_t>(start, (int64_t)cols - (int64_t)b));
  *w0 = (uint64_t)start;
  *B = b;
}

template <typename F>
void parallel_rows(uint32_t rows, F &&fn)
{
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (rows < 65536) nt = 1;
  std::vector<std::thread> th;
  const uint32_t per = (rows + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    const uint32_t a = std::min<uint64_t>((uint64_t)t * per, rows), b = std::min<uint64_t>((uint64_t)(t + 1) * per, rows);
    th.emplace_back([&fn, a, b] { fn(a, b); });
  }
  for (auto &x : th) x.join();
}

} // namespace

extern "C" {

int mfp_generate(const mfp_gen_params *p, uint32_t *rows_out, uint32_t *cols_out, uint64_t *nnz_out,
                 uint32_t *row_ptr, uint32_t *columns, void *values, int dtype)
{
  if (!p || !rows_out || !cols_out || !nnz_out) return MFP_ERR_ARG;
  uint32_t rows = 0, cols = 0;
  if (p->family == MFP_GEN_STENCIL2D) {
    if (p->n == 0 || (uint64_t)p->n * p->n >= (1ull << 32)) return MFP_ERR_ARG;
    rows = cols = p->n * p->n;
  } else if (p->family == MFP_GEN_STENCIL3D) {
    if (p->n == 0 || (uint64_t)p->n * p->n * p->n >= (1ull << 32)) return MFP_ERR_ARG;
    rows = cols = p->n * p->n * p->n;
  } else if (p->family == MFP_GEN_POWERLAW || p->family == MFP_GEN_UNIFORM) {
    if (p->rows == 0) return MFP_ERR_ARG;
    if (p->family == MFP_GEN_POWERLAW && !(p->alpha > 1.0)) return MFP_ERR_ARG;
    rows = cols = p->rows;
  } else {
    return MFP_ERR_ARG;
  }
  *rows_out = rows;
  *cols_out = cols;

  // row lengths
  std::vector<uint32_t> len(rows);
  const uint32_t n = p->n;
  parallel_rows(rows, [&](uint32_t a, uint32_t b) {
    for (uint32_t r = a; r < b; r++) {
      if (p->family == MFP_GEN_STENCIL2D) {
        const uint32_t i = r / n, j = r % n;
        len[r] = 1 + (i > 0) + (i + 1 < n) + (j > 0) + (j + 1 < n);
      } else if (p->family == MFP_GEN_STENCIL3D) {
        const uint32_t j = r % n, i = (r / n) % n, k = r / (n * n);
        auto c = [n](uint32_t v) { return 1u + (v > 0) + (v + 1 < n); };
        len[r] = c(i) * c(j) * c(k);
      } else {
        len[r] = row_len(*p, cols, r);
      }
    }
  });
  uint64_t nnz = 0;
  for (uint32_t r = 0; r < rows; r++) nnz += len[r];
  *nnz_out = nnz;
  if (!row_ptr) return MFP_OK;
  if (nnz >= (1ull << 32)) return MFP_ERR_ARG;
  if (!columns || !values) return MFP_ERR_ARG;
  if (dtype < MFP_F32 || dtype > MFP_F16) return MFP_ERR_ARG;
  row_ptr[0] = 0;
  for (uint32_t r = 0; r < rows; r++) row_ptr[r + 1] = row_ptr[r] + len[r];

  const int vmode = p->values;
  parallel_rows(rows, [&](uint32_t a, uint32_t b) {
    for (uint32_t r = a; r < b; r++) {
      uint64_t e = row_ptr[r];
      if (p->family == MFP_GEN_STENCIL2D) {
        const uint32_t i = r / n, j = r % n;
        const int64_t cand[5] = {i > 0 ? (int64_t)r - n : -1, j > 0 ? (int64_t)r - 1 : -1, (int64_t)r,
                                 j + 1 < n ? (int64_t)r + 1 : -1, i + 1 < n ? (int64_t)r + n : -1};
        for (int64_t c : cand) {
          if (c < 0) continue;
          columns[e] = (uint32_t)c;
          store_value(values, dtype, e, draw_value(vmode, p->seed, e, c == r, 4.0));
          e++;
        }
      } else if (p->family == MFP_GEN_STENCIL3D) {
        const int64_t j = r % n, i = (r / n) % n, k = r / ((uint64_t)n * n);
        for (int dk = -1; dk <= 1; dk++)
          for (int di = -1; di <= 1; di++)
            for (int dj = -1; dj <= 1; dj++) {
              const int64_t kk = k + dk, ii = i + di, jj = j + dj;
              if (kk < 0 || ii < 0 || jj < 0 || kk >= n || ii >= n || jj >= n) continue;
              const uint32_t c = (uint32_t)((kk * n + ii) * n + jj);
              columns[e] = c;
              store_value(values, dtype, e, draw_value(vmode, p->seed, e, c == r, 26.0));
              e++;
            }
      } else {
        const uint32_t L = len[r];
        uint64_t w0, B;
        window(*p, rows, cols, r, L, &w0, &B);
        for (uint32_t t = 0; t < L; t++) {
          const uint64_t lo = w0 + (uint64_t)t * B / L, hi = w0 + ((uint64_t)t + 1) * B / L;
          const uint64_t c = lo + hash3(p->seed, S_COL, e) % (hi - lo);
          columns[e] = (uint32_t)c;
          store_value(values, dtype, e, draw_value(vmode, p->seed, e, c == r, (double)L));
          e++;
        }
      }
    }
  });
  return MFP_OK;
}

int mfp_generate_vector(uint64_t seed, uint64_t n, void *x, int dtype)
{
  if (!x && n) return MFP_ERR_ARG;
  if (dtype < MFP_F32 || dtype > MFP_F16) return MFP_ERR_ARG;
  for (uint64_t i = 0; i < n; i++) store_value(x, dtype, i, u_pm1(hash3(seed, S_VEC, i)));
  return MFP_OK;
}

} // extern "C"
