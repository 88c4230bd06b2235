// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only):
// the seeded generator (csrc/generate.cpp), the MatrixMarket reader of
// include/mfp_compat.hpp and the oracle's C restatement, on small inputs and
// on the malformed and edge cases they must refuse or survive.  Built and run
// by tests/test_host_sanitizers.py; prints "ALL OK" and exits 0 on success.
#include "mfp_compat.hpp"

extern "C" {
#include "../oracle/mfp_oracle.h"
}

#include <cstdio>
#include <cstdlib>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#define REQUIRE(c)                                                                                                     \
  do {                                                                                                                 \
    if (!(c)) {                                                                                                        \
      std::fprintf(stderr, "FAILED %s at line %d\n", #c, __LINE__);                                                    \
      std::exit(1);                                                                                                    \
    }                                                                                                                  \
  } while (0)

namespace {

size_t value_size(int dtype) { return dtype == MFP_F64 ? 8 : dtype == MFP_F32 ? 4 : 2; }

void check_generator(const mfp_gen_params &p, int dtype)
{
  uint32_t rows = 0, cols = 0;
  uint64_t nnz = 0;
  REQUIRE(mfp_generate(&p, &rows, &cols, &nnz, nullptr, nullptr, nullptr, dtype) == MFP_OK);
  std::vector<uint32_t> rp(rows + 1), ci(nnz ? nnz : 1);
  std::vector<unsigned char> va((nnz ? nnz : 1) * value_size(dtype));
  uint32_t rows2 = 0, cols2 = 0;
  uint64_t nnz2 = 0;
  REQUIRE(mfp_generate(&p, &rows2, &cols2, &nnz2, rp.data(), ci.data(), va.data(), dtype) == MFP_OK);
  REQUIRE(rows2 == rows && cols2 == cols && nnz2 == nnz);
  REQUIRE(rp[0] == 0 && rp[rows] == nnz);
  for (uint32_t r = 0; r < rows; r++) {
    REQUIRE(rp[r] <= rp[r + 1]);
    for (uint32_t e = rp[r]; e < rp[r + 1]; e++) {
      REQUIRE(ci[e] < cols);
      if (e > rp[r]) REQUIRE(ci[e - 1] < ci[e]); // ascending, no duplicates
    }
  }
}

void check_generators()
{
  const int dtypes[] = {MFP_F32, MFP_F64, MFP_BF16, MFP_F16};
  std::vector<mfp_gen_params> ps;
  auto add = [&](int fam, int val, uint32_t n, uint32_t rows, uint32_t lo, uint32_t hi, double alpha, double bw) {
    mfp_gen_params p{};
    p.family = fam;
    p.values = val;
    p.seed = 7;
    p.n = n;
    p.rows = rows;
    p.nnz_lo = lo;
    p.nnz_hi = hi;
    p.alpha = alpha;
    p.bandwidth = bw;
    ps.push_back(p);
  };
  add(MFP_GEN_STENCIL2D, MFP_VAL_LAPLACE, 1, 0, 0, 0, 2.0, 1.0);
  add(MFP_GEN_STENCIL2D, MFP_VAL_UNIFORM, 37, 0, 0, 0, 2.0, 1.0);
  add(MFP_GEN_STENCIL3D, MFP_VAL_NORMAL, 2, 0, 0, 0, 2.0, 1.0);
  add(MFP_GEN_STENCIL3D, MFP_VAL_NORMAL, 13, 0, 0, 0, 2.0, 1.0);
  add(MFP_GEN_POWERLAW, MFP_VAL_UNIFORM, 0, 1, 1, 1, 2.0, 1.0);
  add(MFP_GEN_POWERLAW, MFP_VAL_UNIFORM, 0, 3000, 1, 2500, 2.0, 1.0);
  add(MFP_GEN_POWERLAW, MFP_VAL_NORMAL, 0, 3000, 1, 100000, 1.5, 0.01); // rows longer than the window
  add(MFP_GEN_POWERLAW, MFP_VAL_NORMAL, 0, 5000, 2, 64, 3.0, 0.2);
  add(MFP_GEN_UNIFORM, MFP_VAL_UNIFORM, 0, 70000, 4, 64, 2.0, 1.0);  // multi-threaded rows
  add(MFP_GEN_UNIFORM, MFP_VAL_UNIFORM, 0, 10, 20, 40, 2.0, 1.0);     // rows longer than cols
  add(MFP_GEN_UNIFORM, MFP_VAL_NORMAL, 0, 4096, 0, 3, 2.0, 0.001);
  for (const auto &p : ps)
    for (int dt : dtypes) check_generator(p, dt);
  std::vector<float> x(1000);
  REQUIRE(mfp_generate_vector(0, x.size(), x.data(), MFP_F32) == MFP_OK);
  for (float v : x) REQUIRE(v >= -1.0f && v < 1.0f);
  mfp_gen_params bad{};
  bad.family = 99;
  uint32_t r = 0, c = 0;
  uint64_t n = 0;
  REQUIRE(mfp_generate(&bad, &r, &c, &n, nullptr, nullptr, nullptr, MFP_F32) != MFP_OK);
  REQUIRE(mfp_generate(nullptr, &r, &c, &n, nullptr, nullptr, nullptr, MFP_F32) != MFP_OK);
}

bool reads(const std::string &text, matrix_market::matrix_class *out = nullptr)
{
  std::istringstream is(text);
  try {
    matrix_market::reader rd(is);
    if (out) *out = rd.matrix();
    return true;
  } catch (const std::runtime_error &) {
    return false;
  }
}

void check_reader()
{
  matrix_market::matrix_class m;
  REQUIRE(reads("%%MatrixMarket matrix coordinate real general\n% c\n\n3 4 3\n1 1 1.5\n3 4 -2\n2 2 0.25\n", &m));
  REQUIRE(m.meta.rows_count == 3 && m.meta.cols_count == 4 && m.meta.non_zero_count == 3);
  REQUIRE(m.get_row_ids()[1] == 2 && m.get_col_ids()[1] == 3 && m.get_dbl_data()[1] == -2.0);
  REQUIRE(reads("%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 1\n2 1\n", &m));
  REQUIRE(m.get_dbl_data()[0] == 1.0);
  REQUIRE(reads("%%MatrixMarket matrix coordinate integer symmetric\n2 2 1\n2 1 7\n", &m));
  REQUIRE(m.meta.matrix_storage_scheme == matrix_market::matrix_class::storage_scheme::symmetric);
  REQUIRE(reads("%%MatrixMarket matrix coordinate real general\n0 0 0\n", &m));
  const char *bad[] = {
      "",
      "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
      "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
      "%%MatrixMarket matrix coordinate real banded\n1 1 1\n1 1 1\n",
      "%%MatrixMarket matrix coordinate real general\n",
      "%%MatrixMarket matrix coordinate real general\nx y z\n",
      "%%MatrixMarket matrix coordinate real general\n2 2 1\n1\n",
      "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",
      "%%MatrixMarket matrix coordinate real general\n2 2 1\n0 1 1.0\n",
      "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n",
      "%%MatrixMarket matrix coordinate real general\n8589934592 2 1\n1 1 1.0\n",
      "not a banner\n1 1 1\n1 1 1\n",
  };
  for (const char *t : bad) REQUIRE(!reads(t));
}

void check_oracle()
{
  std::mt19937 g(5);
  const uint32_t rows = 300, cols = 200;
  std::vector<uint32_t> tr, tc;
  std::vector<double> tv;
  for (uint32_t r = 0; r < rows; r++) {
    const uint32_t len = g() % 40;
    for (uint32_t k = 0; k < len; k++) {
      tr.push_back(r);
      tc.push_back(g() % cols);
      tv.push_back((double)(g() % 2001) / 1000.0 - 1.0);
    }
  }
  const uint64_t nnz = tr.size();
  for (int dt : {ORC_F32, ORC_F64}) {
    const size_t s = dt == ORC_F32 ? 4 : 8;
    std::vector<uint32_t> rp(rows + 1), ci(nnz);
    std::vector<unsigned char> va(nnz * s), x(cols * s), y(rows * s), y2(rows * s);
    REQUIRE(orc_csr_from_triplets(rows, cols, nnz, tr.data(), tc.data(), tv.data(), dt, rp.data(), ci.data(),
                                  va.data()) == 0);
    for (uint32_t j = 0; j < cols; j++) {
      if (dt == ORC_F32) ((float *)x.data())[j] = 0.5f;
      else ((double *)x.data())[j] = 0.5;
    }
    uint32_t st[3];
    double sd = 0;
    orc_rows_statistics(rows, rp.data(), st, &sd);
    const uint32_t w = st[1];
    std::vector<unsigned char> ed((size_t)w * rows * s + s);
    std::vector<uint32_t> ec((size_t)w * rows + 1);
    orc_ell_from_csr(rows, rp.data(), ci.data(), va.data(), dt, w, ed.data(), ec.data());
    orc_spmv_ell_seq(dt, rows, w, ec.data(), ed.data(), x.data(), y.data());
    orc_spmv_csr_seq(dt, rows, rp.data(), ci.data(), va.data(), x.data(), y2.data());
    REQUIRE(y == y2);
    const uint32_t hw = orc_hybrid_width(rows, rp.data(), 0.2);
    const uint64_t nc = orc_coo_count(rows, rp.data(), hw);
    std::vector<unsigned char> cd(nc * s + s);
    std::vector<uint32_t> cc(nc + 1), cr(nc + 1);
    orc_coo_from_csr(rows, rp.data(), ci.data(), va.data(), dt, hw, cd.data(), cc.data(), cr.data());
    orc_spmv_coo_seq(dt, nc, cr.data(), cc.data(), cd.data(), x.data(), y.data());
    for (uint32_t lanes = 2; lanes <= 64; lanes *= 2)
      orc_spmv_csr_lanes(dt, lanes, rows, rp.data(), ci.data(), va.data(), x.data(), y.data());
    const uint32_t nb = orc_fill_row_blocks(0, rows, rp.data(), nullptr, 64);
    std::vector<uint32_t> blocks(nb + 1);
    REQUIRE(orc_fill_row_blocks(1, rows, rp.data(), blocks.data(), 64) == nb);
    orc_spmv_csr_adaptive(dt, rows, rp.data(), ci.data(), va.data(), x.data(), y.data(), 64, 256, 16, 8);
    const uint32_t sc = orc_scoo_slices_count(rows, 4, 16384, dt);
    const uint32_t ss = (rows + sc - 1) / sc;
    std::vector<unsigned char> sv(nnz * s + s);
    std::vector<uint32_t> sci(nnz + 1), sri(nnz + 1), sp(sc + 1);
    orc_scoo_build(rows, rp.data(), ci.data(), va.data(), dt, sc, ss, sv.data(), sci.data(), sri.data(), sp.data());
    std::vector<double> y64(rows), ad(rows);
    orc_spmv_csr_f64(dt, rows, rp.data(), ci.data(), va.data(), x.data(), y64.data(), ad.data());
    REQUIRE(orc_spmv_csr_multi_thread(dt, 3, rows, rp.data(), ci.data(), va.data(), x.data(), y.data()) >= 0.0);
    REQUIRE(orc_compare_error(dt, rows, y.data(), y.data()) == 0.0);
  }
  std::vector<uint32_t> badr = {0, 5}, badc = {0, 0};
  std::vector<double> badv = {1.0, 1.0};
  std::vector<uint32_t> rp(3), ci(2);
  std::vector<float> va(2);
  REQUIRE(orc_csr_from_triplets(2, 2, 2, badr.data(), badc.data(), badv.data(), ORC_F32, rp.data(), ci.data(),
                                va.data()) == -1);
}

} // namespace

int main()
{
  check_generators();
  check_reader();
  check_oracle();
  std::printf("ALL OK\n");
  return 0;
}
