// Format selector network: MLP 64 -> 128 -> 128 -> 8 (ReLU, softmax), 25,864
// FP32 parameters, trained on the GPU with these kernels:
//   forward    H1 = relu(X W1 + b1), H2 = relu(H1 W2 + b2), Z = H2 W3 + b3
//   loss       softmax cross-entropy against target distributions P
//   backward   dZ = (softmax(Z) - P) / n; dW3 = H2^T dZ; dH2 = dZ W3^T * [H2 > 0]; ...
//   update     Adam (Kingma & Ba, 2015)
// Inference can run in BF16 on the matrix cores (v_mfma_f32_16x16x16_bf16,
// FP32 accumulation): activations and weights rounded to BF16, bias, ReLU and
// softmax in FP32.
// Weights are row-major W[in][out].
#include "mfp_internal.h"
#include "wave64.h"

#include <hip/hip_bf16.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace mfp {

constexpr int NI = 64, NH = 128, NO = 8;
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + NI * NH, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
                 OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO, NPARAM = OFF_B3 + NO;
static_assert(NPARAM == 25864, "parameter count");

// C[M x N] = alpha * op(A) op(B) + beta * C; op(A) is M x K, op(B) K x N.
// A is M x K (lda K) or, transposed, K x M (lda M); likewise B.
constexpr int GT = 16;
__global__ __launch_bounds__(GT *GT) void k_gemm(const float *__restrict__ A, const float *__restrict__ B,
                                                 float *__restrict__ Cm, int M, int N, int K, int ta, int tb,
                                                 float alpha, float beta)
{
  __shared__ float sa[GT][GT + 1];
  __shared__ float sb[GT][GT + 1];
  const int tx = threadIdx.x % GT, ty = threadIdx.x / GT;
  const int row = blockIdx.y * GT + ty, col = blockIdx.x * GT + tx;
  float acc = 0.f;
  for (int k0 = 0; k0 < K; k0 += GT) {
    const int ka = k0 + tx, kb = k0 + ty;
    sa[ty][tx] = (row < M && ka < K) ? (ta ? A[(size_t)ka * M + row] : A[(size_t)row * K + ka]) : 0.f;
    sb[ty][tx] = (kb < K && col < N) ? (tb ? B[(size_t)col * K + kb] : B[(size_t)kb * N + col]) : 0.f;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GT; k++) acc = __builtin_fmaf(sa[ty][k], sb[k][tx], acc);
    __syncthreads();
  }
  if (row < M && col < N) {
    float *c = Cm + (size_t)row * N + col;
    *c = beta == 0.f ? alpha * acc : alpha * acc + beta * *c;
  }
}

__global__ void k_bias_act(float *__restrict__ H, const float *__restrict__ b, int M, int N, int relu)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  float v = H[i] + b[i % N];
  H[i] = relu ? fmaxf(v, 0.f) : v;
}

// per row: softmax, loss -sum p log q, dZ = (q - p) / n
__global__ void k_softmax_xent(const float *__restrict__ Z, const float *__restrict__ P, int M,
                               float *__restrict__ dZ, float *__restrict__ Q, double *__restrict__ loss)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= M) return;
  float z[NO], mx = -INFINITY;
  for (int j = 0; j < NO; j++) {
    z[j] = Z[r * NO + j];
    mx = fmaxf(mx, z[j]);
  }
  float s = 0.f;
  for (int j = 0; j < NO; j++) s += expf(z[j] - mx);
  const float ls = logf(s);
  double l = 0.0;
  for (int j = 0; j < NO; j++) {
    const float lq = z[j] - mx - ls;
    const float q = expf(lq);
    if (Q) Q[r * NO + j] = q;
    if (P) {
      const float p = P[r * NO + j];
      if (p > 0.f) l -= (double)p * (double)lq;
      if (dZ) dZ[r * NO + j] = (q - p) / (float)M;
    }
  }
  if (loss && P) atomicAdd(loss, l / M);
}

__global__ void k_relu_bwd(float *__restrict__ dH, const float *__restrict__ H, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && H[i] <= 0.f) dH[i] = 0.f;
}

__global__ void k_colsum(const float *__restrict__ D, int M, int N, float *__restrict__ out)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int i = 0; i < M; i++) s += D[(size_t)i * N + j];
  out[j] = s;
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1,
                       float c2)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- BF16 inference on MFMA -------------------------------------------------------
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float f)
{
  __hip_bfloat16 h = __float2bfloat16(f);
  return *reinterpret_cast<short *>(&h);
}

// C[M x N] = A[M x K] W[K x N] (+ bias, ReLU) with A, W rounded to BF16.
// One wave per 16 x 16 output tile; K in steps of 16.  Operand layout of
// v_mfma_f32_16x16x16_bf16: lane l holds A[l % 16][4 (l / 16) + j] and
// W[4 (l / 16) + j][l % 16] (j = 0..3); the result D[4 (l / 16) + j][l % 16].
__global__ __launch_bounds__(64) void k_mfma_linear(const float *__restrict__ A, const float *__restrict__ W,
                                                    const float *__restrict__ bias, float *__restrict__ C,
                                                    int M, int N, int K, int relu)
{
  const int lane = threadIdx.x;
  const int tm = blockIdx.y * 16, tn = blockIdx.x * 16;
  const int r = lane % 16, kq = 4 * (lane / 16);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int k = k0 + kq + j;
      const int ar = tm + r, bc = tn + r;
      a[j] = to_bf16((ar < M && k < K) ? A[(size_t)ar * K + k] : 0.f);
      b[j] = to_bf16((k < K && bc < N) ? W[(size_t)k * N + bc] : 0.f);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int row = tm + kq + j, col = tn + r;
    if (row < M && col < N) {
      float v = acc[j] + bias[col];
      C[(size_t)row * N + col] = relu ? fmaxf(v, 0.f) : v;
    }
  }
}

} // namespace mfp

using namespace mfp;

struct mfp_mlp_s {
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
  float *X = nullptr, *P = nullptr, *H1 = nullptr, *H2 = nullptr, *Z = nullptr, *Q = nullptr;
  float *dZ = nullptr, *dH2 = nullptr, *dH1 = nullptr;
  double *loss = nullptr;
  uint32_t cap = 0;
  uint64_t step = 0;
};

static void mlp_free_batch(mfp_mlp m)
{
  for (float **p : {&m->X, &m->P, &m->H1, &m->H2, &m->Z, &m->Q, &m->dZ, &m->dH2, &m->dH1}) {
    dfree(*p);
    *p = nullptr;
  }
  m->cap = 0;
}

static int mlp_reserve(mfp_mlp m, uint32_t n)
{
  if (n <= m->cap) return MFP_OK;
  mlp_free_batch(m);
  const uint64_t c = n;
  MFP_CHECK(dmalloc((void **)&m->X, c * NI * 4));
  MFP_CHECK(dmalloc((void **)&m->P, c * NO * 4));
  MFP_CHECK(dmalloc((void **)&m->H1, c * NH * 4));
  MFP_CHECK(dmalloc((void **)&m->H2, c * NH * 4));
  MFP_CHECK(dmalloc((void **)&m->Z, c * NO * 4));
  MFP_CHECK(dmalloc((void **)&m->Q, c * NO * 4));
  MFP_CHECK(dmalloc((void **)&m->dZ, c * NO * 4));
  MFP_CHECK(dmalloc((void **)&m->dH2, c * NH * 4));
  MFP_CHECK(dmalloc((void **)&m->dH1, c * NH * 4));
  m->cap = n;
  return MFP_OK;
}

static void gemm(const float *A, const float *B, float *C, int M, int N, int K, int ta, int tb, float beta = 0.f)
{
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
  k_gemm<<<grid, GT * GT>>>(A, B, C, M, N, K, ta, tb, 1.f, beta);
}

static int mlp_forward(mfp_mlp m, uint32_t n)
{
  const float *p = m->params;
  gemm(m->X, p + OFF_W1, m->H1, n, NH, NI, 0, 0);
  k_bias_act<<<ceil_div_u32((uint64_t)n * NH, 256), 256>>>(m->H1, p + OFF_B1, n, NH, 1);
  gemm(m->H1, p + OFF_W2, m->H2, n, NH, NH, 0, 0);
  k_bias_act<<<ceil_div_u32((uint64_t)n * NH, 256), 256>>>(m->H2, p + OFF_B2, n, NH, 1);
  gemm(m->H2, p + OFF_W3, m->Z, n, NO, NH, 0, 0);
  k_bias_act<<<ceil_div_u32((uint64_t)n * NO, 256), 256>>>(m->Z, p + OFF_B3, n, NO, 0);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

static int mlp_backward(mfp_mlp m, uint32_t n)
{
  const float *p = m->params;
  float *g = m->grads;
  MFP_CHECK(fill_bytes_zero(m->loss, 8, nullptr));
  k_softmax_xent<<<ceil_div_u32(n, 256), 256>>>(m->Z, m->P, n, m->dZ, m->Q, m->loss);
  gemm(m->H2, m->dZ, g + OFF_W3, NH, NO, n, 1, 0);                       // dW3 = H2^T dZ
  k_colsum<<<1, 64>>>(m->dZ, n, NO, g + OFF_B3);                           // db3
  gemm(m->dZ, p + OFF_W3, m->dH2, n, NH, NO, 0, 1);                       // dH2 = dZ W3^T
  k_relu_bwd<<<ceil_div_u32((uint64_t)n * NH, 256), 256>>>(m->dH2, m->H2, n * NH);
  gemm(m->H1, m->dH2, g + OFF_W2, NH, NH, n, 1, 0);                       // dW2 = H1^T dH2
  k_colsum<<<1, NH>>>(m->dH2, n, NH, g + OFF_B2);
  gemm(m->dH2, p + OFF_W2, m->dH1, n, NH, NH, 0, 1);                      // dH1 = dH2 W2^T
  k_relu_bwd<<<ceil_div_u32((uint64_t)n * NH, 256), 256>>>(m->dH1, m->H1, n * NH);
  gemm(m->X, m->dH1, g + OFF_W1, NI, NH, n, 1, 0);                        // dW1 = X^T dH1
  k_colsum<<<1, NH>>>(m->dH1, n, NH, g + OFF_B1);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

extern "C" {

int mfp_mlp_create(mfp_mlp *out, uint64_t seed)
{
  MFP_ARG(out, "null pointer");
  MFP_CHECK(ensure_init());
  auto *m = new mfp_mlp_s;
  int st = dmalloc((void **)&m->params, NPARAM * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->grads, NPARAM * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->m1, NPARAM * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->m2, NPARAM * 4);
  if (st == MFP_OK) st = dmalloc((void **)&m->loss, 8);
  if (st == MFP_OK) st = fill_bytes_zero(m->m1, NPARAM * 4, nullptr);
  if (st == MFP_OK) st = fill_bytes_zero(m->m2, NPARAM * 4, nullptr);
  if (st != MFP_OK) {
    mfp_mlp_destroy(m);
    return st;
  }
  // He-uniform weights from a splitmix64 stream, zero biases
  std::vector<float> h(NPARAM, 0.f);
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t v = z;
    v = (v ^ (v >> 30)) * 0xBF58476D1CE4E5B9ull;
    v = (v ^ (v >> 27)) * 0x94D049BB133111EBull;
    v ^= v >> 31;
    return ((double)(v >> 11) + 0.5) / 9007199254740992.0 * 2.0 - 1.0;
  };
  auto init = [&](size_t off, int fan_in, int cnt) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < cnt; i++) h[off + i] = (float)(a * next());
  };
  init(OFF_W1, NI, NI * NH);
  init(OFF_W2, NH, NH * NH);
  init(OFF_W3, NH, NH * NO);
  st = mfp_memcpy_h2d(m->params, h.data(), NPARAM * 4);
  if (st != MFP_OK) {
    mfp_mlp_destroy(m);
    return st;
  }
  *out = m;
  return MFP_OK;
}

int mfp_mlp_destroy(mfp_mlp m)
{
  if (!m) return MFP_OK;
  mlp_free_batch(m);
  dfree(m->params);
  dfree(m->grads);
  dfree(m->m1);
  dfree(m->m2);
  dfree(m->loss);
  delete m;
  return MFP_OK;
}

uint32_t mfp_mlp_param_count(void) { return (uint32_t)NPARAM; }

int mfp_mlp_get_params(mfp_mlp m, float *out)
{
  MFP_ARG(m && out, "null pointer");
  return mfp_memcpy_d2h(out, m->params, NPARAM * 4);
}

int mfp_mlp_set_params(mfp_mlp m, const float *in)
{
  MFP_ARG(m && in, "null pointer");
  m->step = 0;
  MFP_CHECK(fill_bytes_zero(m->m1, NPARAM * 4, nullptr));
  MFP_CHECK(fill_bytes_zero(m->m2, NPARAM * 4, nullptr));
  return mfp_memcpy_h2d(m->params, in, NPARAM * 4);
}

int mfp_mlp_loss_grad(mfp_mlp m, const float *X, const float *P, uint32_t n, double *loss, float *grad)
{
  MFP_ARG(m && X && P && n > 0, "null pointer or empty batch");
  MFP_CHECK(mlp_reserve(m, n));
  MFP_CHECK(mfp_memcpy_h2d(m->X, X, (size_t)n * NI * 4));
  MFP_CHECK(mfp_memcpy_h2d(m->P, P, (size_t)n * NO * 4));
  MFP_CHECK(mlp_forward(m, n));
  MFP_CHECK(mlp_backward(m, n));
  if (loss) MFP_CHECK(mfp_memcpy_d2h(loss, m->loss, 8));
  if (grad) MFP_CHECK(mfp_memcpy_d2h(grad, m->grads, NPARAM * 4));
  return MFP_OK;
}

int mfp_mlp_train_step(mfp_mlp m, const float *X, const float *P, uint32_t n, float lr, double *loss)
{
  MFP_CHECK(mfp_mlp_loss_grad(m, X, P, n, loss, nullptr));
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
  k_adam<<<ceil_div_u32(NPARAM, 256), 256>>>(m->params, m->grads, m->m1, m->m2, (int)NPARAM, lr, b1, b2,
                                             1e-8f, c1, c2);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

int mfp_mlp_predict(mfp_mlp m, const float *X, uint32_t n, int bf16, float *probs)
{
  MFP_ARG(m && X && probs && n > 0, "null pointer or empty batch");
  MFP_CHECK(mlp_reserve(m, n));
  MFP_CHECK(mfp_memcpy_h2d(m->X, X, (size_t)n * NI * 4));
  if (bf16) {
    const float *p = m->params;
    const dim3 g1(NH / 16, (n + 15) / 16), g3((NO + 15) / 16, (n + 15) / 16);
    k_mfma_linear<<<g1, 64>>>(m->X, p + OFF_W1, p + OFF_B1, m->H1, n, NH, NI, 1);
    k_mfma_linear<<<g1, 64>>>(m->H1, p + OFF_W2, p + OFF_B2, m->H2, n, NH, NH, 1);
    k_mfma_linear<<<g3, 64>>>(m->H2, p + OFF_W3, p + OFF_B3, m->Z, n, NO, NH, 0);
    MFP_HIP(hipGetLastError());
  } else {
    MFP_CHECK(mlp_forward(m, n));
  }
  k_softmax_xent<<<ceil_div_u32(n, 256), 256>>>(m->Z, nullptr, n, nullptr, m->Q, nullptr);
  MFP_HIP(hipGetLastError());
  return mfp_memcpy_d2h(probs, m->Q, (size_t)n * NO * 4);
}

} // extern "C"
