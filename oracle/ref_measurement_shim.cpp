// C entry points over the reference's own measurement_class sources
// (common/measurement_class.{h,cpp}), compiled where they lie.  TEST
// INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's measurement
// arithmetic and the product's compare_results text.
#include "measurement_class.h"

#include <cstring>
#include <sstream>
#include <string>

extern "C" {

// measurement_class ctor (measurement_class.cpp:7-16); out2 = {GB/s, GFLOP/s}
void ref_measurement(const char *format, double elapsed, double bytes,
                     double ops, double *out2)
{
  measurement_class m(format, elapsed, bytes, ops);
  out2[0] = m.get_effective_bandwidth();
  out2[1] = m.get_computational_throughput();
}

// Mean of n records via operator+= and finalize (measurement_class.h:29-49);
// out3 = {elapsed, GB/s, GFLOP/s}
void ref_measurement_mean(unsigned n, const double *elapsed,
                          const double *bytes, const double *ops,
                          double *out3)
{
  measurement_class acc;
  for (unsigned i = 0; i < n; i++)
    acc += measurement_class("x", elapsed[i], bytes[i], ops[i]);
  acc.finalize();
  out3[0] = acc.get_elapsed();
  out3[1] = acc.get_effective_bandwidth();
  out3[2] = acc.get_computational_throughput();
}

// compare_results<T> (measurement_class.h:60-89) with std::cerr captured.
static void capture(std::string &dst, void (*fn)(unsigned, const void *, const void *),
                    unsigned n, const void *a, const void *b)
{
  std::ostringstream os;
  auto *old = std::cerr.rdbuf(os.rdbuf());
  fn(n, a, b);
  std::cerr.rdbuf(old);
  dst = os.str();
}

static void cmp_f32(unsigned n, const void *a, const void *b)
{
  compare_results<float>(n, static_cast<const float *>(a), static_cast<const float *>(b));
}
static void cmp_f64(unsigned n, const void *a, const void *b)
{
  compare_results<double>(n, static_cast<const double *>(a), static_cast<const double *>(b));
}

size_t ref_compare_results(int dtype, unsigned n, const void *a, const void *b,
                           char *out, size_t cap)
{
  std::string s;
  capture(s, dtype == 0 ? cmp_f32 : cmp_f64, n, a, b);
  if (cap) {
    size_t k = s.size() < cap - 1 ? s.size() : cap - 1;
    std::memcpy(out, s.data(), k);
    out[k] = 0;
  }
  return s.size();
}
}
