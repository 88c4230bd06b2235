// Runtime: device selection, memory, grow-only buffers, timers, fills and the
// Infinity-Cache flush.
#include "mfp_internal.h"
#include "wave64.h"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstring>
#include <mutex>

namespace mfp {

static thread_local std::string g_last_error;
thread_local kernel_marks *g_kernel_marks = nullptr;
void set_error(const std::string &msg) { g_last_error = msg; }
const std::string &last_error() { return g_last_error; }

static device_props g_props;
static std::mutex g_init_mutex;

const device_props &props() { return g_props; }

static int init_device(int device)
{
  std::lock_guard<std::mutex> lock(g_init_mutex);
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    set_error("no HIP device visible");
    return MFP_ERR_NODEV;
  }
  if (device < 0 || device >= count) {
    set_error("device index out of range");
    return MFP_ERR_ARG;
  }
  MFP_HIP(hipSetDevice(device));
  hipDeviceProp_t p;
  MFP_HIP(hipGetDeviceProperties(&p, device));
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("libmfp is built for gfx950, found ") + p.gcnArchName);
    return MFP_ERR_NODEV;
  }
  g_props.device = device;
  g_props.cu_count = p.multiProcessorCount;
  g_props.lds_per_block = p.sharedMemPerBlock;
  std::strncpy(g_props.name, p.name, sizeof(g_props.name) - 1);
  std::strncpy(g_props.arch, p.gcnArchName, sizeof(g_props.arch) - 1);
  g_props.xcd_count = 8;
  size_t fr = 0, tot = 0;
  MFP_HIP(hipMemGetInfo(&fr, &tot));
  g_props.free_mem = fr;
  g_props.total_mem = tot;
  // the device address of the static scratch of mfp_plan_spmv_noalloc
  return static_scratch_setup();
}

int ensure_init()
{
  if (g_props.device >= 0) {
    int cur = -1;
    MFP_HIP(hipGetDevice(&cur));
    if (cur != g_props.device) MFP_HIP(hipSetDevice(g_props.device));
    return MFP_OK;
  }
  return init_device(0);
}

// Every allocation carries 256 bytes of slack past its end: kernels that read
// 16-byte vectors may touch up to 12 bytes beyond an array's last element.
int dmalloc(void **p, size_t bytes)
{
  MFP_CHECK(ensure_init());
  *p = nullptr;
  MFP_HIP(hipMalloc(p, bytes + 256));
  return MFP_OK;
}

void dfree(void *p)
{
  if (p) (void)hipFree(p);
}

// ---- fills ----------------------------------------------------------------

template <typename T>
__global__ void k_fill(T *__restrict__ p, uint64_t n, T v)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename T>
static int launch_fill(T *p, uint64_t n, T v, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  k_fill<T><<<(uint32_t)blocks, 256, 0, s>>>(p, n, v);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

int fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) { return launch_fill(p, n, v, s); }

int fill_bytes_zero(void *p, size_t bytes, hipStream_t s)
{
  // Zeroing with a kernel (not hipMemsetAsync) keeps captured graphs replayable.
  if (bytes % 4 == 0) return launch_fill((uint32_t *)p, bytes / 4, 0u, s);
  return launch_fill((uint8_t *)p, bytes, (uint8_t)0, s);
}

template <typename TI, typename TO>
__global__ void k_convert(const TI *__restrict__ in, TO *__restrict__ out, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (TO)(float)in[i];
}
template <typename TI>
__global__ void k_convert_exact(const TI *__restrict__ in, double *__restrict__ out, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (double)in[i];
}
__global__ void k_convert_d2f(const double *__restrict__ in, float *__restrict__ out, uint64_t n)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (float)in[i];
}

int convert_values(int din, const void *in, int dout, void *out, uint64_t n, hipStream_t s)
{
  if (n == 0) return MFP_OK;
  uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 65536);
  if (din == dout) {
    MFP_HIP(hipMemcpyAsync(out, in, n * dtype_size(din), hipMemcpyDeviceToDevice, s));
    return MFP_OK;
  }
  if (din == MFP_F64 && dout == MFP_F32) {
    k_convert_d2f<<<blocks, 256, 0, s>>>((const double *)in, (float *)out, n);
  } else if (dout == MFP_F64 && din == MFP_F32) {
    k_convert_exact<float><<<blocks, 256, 0, s>>>((const float *)in, (double *)out, n);
  } else {
    set_error("unsupported value conversion");
    return MFP_ERR_ARG;
  }
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

// STREAM-style copy with 16-byte accesses: the measured HBM reference point.
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const u32x4_t *__restrict__ src, u32x4_t *__restrict__ dst,
                                                uint64_t n16)
{
  // four independent 16-byte loads in flight per lane
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const u32x4_t a = __builtin_nontemporal_load(src + i);
    const u32x4_t b = __builtin_nontemporal_load(src + i + stride);
    const u32x4_t c = __builtin_nontemporal_load(src + i + 2 * stride);
    const u32x4_t d = __builtin_nontemporal_load(src + i + 3 * stride);
    __builtin_nontemporal_store(a, dst + i);
    __builtin_nontemporal_store(b, dst + i + stride);
    __builtin_nontemporal_store(c, dst + i + 2 * stride);
    __builtin_nontemporal_store(d, dst + i + 3 * stride);
  }
  for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Cache flush: a read of `bytes` of scratch (written once, when it is
// allocated).  Reading replaces the Infinity Cache's lines with clean ones, so
// the timed kernel that follows pays for no write-back of the flush itself;
// a write flush would leave up to 256 MB of dirty lines behind.
static void *g_flush_buf = nullptr;
static size_t g_flush_bytes = 0;

__global__ void k_read_flush(const u32x4_t *__restrict__ p, uint64_t n16, uint32_t *__restrict__ sink)
{
  uint32_t acc = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x) {
    const u32x4_t v = p[i];
    acc ^= v[0] ^ v[1] ^ v[2] ^ v[3];
  }
  // the fill pattern XORs to zero, so this store never happens; it keeps the loads
  if (acc == 0xFFFFFFFFu) sink[0] = acc;
}

} // namespace mfp

using namespace mfp;

// ---- C ABI -------------------------------------------------------------------

extern "C" {

int mfp_init(int device)
{
  if (g_props.device == device) return ensure_init();
  g_props.device = -1;
  return init_device(device);
}

int mfp_device_query(mfp_device_info *out)
{
  MFP_ARG(out, "out is null");
  MFP_CHECK(ensure_init());
  size_t fr = 0, tot = 0;
  MFP_HIP(hipMemGetInfo(&fr, &tot));
  std::memset(out, 0, sizeof(*out));
  out->device = g_props.device;
  out->cu_count = g_props.cu_count;
  out->xcd_count = g_props.xcd_count;
  out->lds_per_block = g_props.lds_per_block;
  out->free_mem = fr;
  out->total_mem = tot;
  std::strncpy(out->name, g_props.name, sizeof(out->name) - 1);
  std::strncpy(out->arch, g_props.arch, sizeof(out->arch) - 1);
  return MFP_OK;
}

const char *mfp_status_string(int status)
{
  switch (status) {
  case MFP_OK: return "ok";
  case MFP_ERR_HIP: return "HIP runtime error";
  case MFP_ERR_ARG: return "invalid argument";
  case MFP_ERR_BUDGET: return "layout exceeds the memory budget";
  case MFP_ERR_STORAGE: return "only general matrices are supported";
  case MFP_ERR_NODEV: return "no gfx950 device";
  case MFP_ERR_STATE: return "invalid state";
  default: return "unknown status";
  }
}

const char *mfp_last_error(void) { return g_last_error.c_str(); }

int mfp_synchronize(void)
{
  MFP_CHECK(ensure_init());
  MFP_HIP(hipDeviceSynchronize());
  return MFP_OK;
}

int mfp_malloc(void **p, size_t bytes)
{
  MFP_ARG(p, "ptr is null");
  return dmalloc(p, bytes);
}

int mfp_free(void *p)
{
  if (!p) return MFP_OK;
  MFP_HIP(hipFree(p));
  return MFP_OK;
}

int mfp_memcpy_h2d(void *dst, const void *src, size_t bytes)
{
  if (bytes == 0) return MFP_OK;
  MFP_CHECK(ensure_init());
  MFP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return MFP_OK;
}

int mfp_memcpy_d2h(void *dst, const void *src, size_t bytes)
{
  if (bytes == 0) return MFP_OK;
  MFP_CHECK(ensure_init());
  MFP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return MFP_OK;
}

int mfp_memcpy_d2d(void *dst, const void *src, size_t bytes, mfp_stream s)
{
  if (bytes == 0) return MFP_OK;
  MFP_CHECK(ensure_init());
  MFP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(s)));
  return MFP_OK;
}

int mfp_fill(int dtype, void *dst, uint64_t n, double value, mfp_stream s)
{
  MFP_CHECK(ensure_init());
  hipStream_t st = as_stream(s);
  switch (dtype) {
  case MFP_F32: return launch_fill((float *)dst, n, (float)value, st);
  case MFP_F64: return launch_fill((double *)dst, n, value, st);
  case MFP_BF16: {
    __hip_bfloat16 b = __float2bfloat16((float)value);
    uint16_t bits;
    std::memcpy(&bits, &b, 2);
    return launch_fill((uint16_t *)dst, n, bits, st);
  }
  case MFP_F16: {
    __half h = __float2half((float)value);
    uint16_t bits;
    std::memcpy(&bits, &h, 2);
    return launch_fill((uint16_t *)dst, n, bits, st);
  }
  default: MFP_ARG(false, "dtype");
  }
}

int mfp_flush_cache(uint64_t bytes, mfp_stream s)
{
  MFP_CHECK(ensure_init());
  if (bytes < (64ull << 20)) bytes = 64ull << 20;
  bytes = (bytes + 15) / 16 * 16;
  if (g_flush_bytes < bytes) {
    dfree(g_flush_buf);
    g_flush_buf = nullptr;
    g_flush_bytes = 0;
    MFP_CHECK(dmalloc(&g_flush_buf, bytes));
    MFP_CHECK(launch_fill((uint32_t *)g_flush_buf, bytes / 4, 0x5A5A5A5Au, as_stream(s)));
    g_flush_bytes = bytes;
  }
  k_read_flush<<<(uint32_t)std::max(1, g_props.cu_count) * 8, 256, 0, as_stream(s)>>>(
      (const u32x4_t *)g_flush_buf, bytes / 16, (uint32_t *)g_flush_buf);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

int mfp_stream_copy(void *dst, const void *src, uint64_t bytes, mfp_stream s)
{
  MFP_CHECK(ensure_init());
  MFP_ARG(bytes % 16 == 0, "bytes must be a multiple of 16");
  MFP_ARG(((uintptr_t)dst | (uintptr_t)src) % 16 == 0, "pointers must be 16-byte aligned");
  const uint64_t n16 = bytes / 16;
  if (n16 == 0) return MFP_OK;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((n16 + 255) / 256, (uint64_t)g_props.cu_count * 32);
  k_copy16<<<grid, 256, 0, as_stream(s)>>>((const u32x4_t *)src, (u32x4_t *)dst, n16);
  MFP_HIP(hipGetLastError());
  return MFP_OK;
}

struct mfp_timer_s {
  hipEvent_t a = nullptr, b = nullptr;
};

int mfp_timer_create(mfp_timer *t)
{
  MFP_ARG(t, "timer is null");
  MFP_CHECK(ensure_init());
  auto *tm = new mfp_timer_s;
  if (hipEventCreate(&tm->a) != hipSuccess || hipEventCreate(&tm->b) != hipSuccess) {
    delete tm;
    set_error("hipEventCreate failed");
    return MFP_ERR_HIP;
  }
  *t = tm;
  return MFP_OK;
}

int mfp_timer_start(mfp_timer t, mfp_stream s)
{
  MFP_HIP(hipEventRecord(t->a, as_stream(s)));
  return MFP_OK;
}

int mfp_timer_stop(mfp_timer t, mfp_stream s)
{
  MFP_HIP(hipEventRecord(t->b, as_stream(s)));
  return MFP_OK;
}

int mfp_timer_elapsed_ms(mfp_timer t, float *ms)
{
  MFP_HIP(hipEventSynchronize(t->b));
  MFP_HIP(hipEventElapsedTime(ms, t->a, t->b));
  return MFP_OK;
}

int mfp_timer_destroy(mfp_timer t)
{
  if (!t) return MFP_OK;
  (void)hipEventDestroy(t->a);
  (void)hipEventDestroy(t->b);
  delete t;
  return MFP_OK;
}

struct mfp_buffer_s {
  void *data = nullptr;
  size_t size = 0;
};

int mfp_buffer_create(mfp_buffer *b)
{
  MFP_ARG(b, "buffer is null");
  *b = new mfp_buffer_s;
  return MFP_OK;
}

int mfp_buffer_resize(mfp_buffer b, size_t bytes)
{
  MFP_ARG(b, "buffer is null");
  if (bytes <= b->size) return MFP_OK;
  dfree(b->data); // resizable_gpu_memory.cpp:16-20: free, then allocate
  b->data = nullptr;
  b->size = 0;
  MFP_CHECK(dmalloc(&b->data, bytes));
  b->size = bytes;
  return MFP_OK;
}

int mfp_buffer_clear(mfp_buffer b)
{
  MFP_ARG(b, "buffer is null");
  dfree(b->data);
  b->data = nullptr;
  b->size = 0;
  return MFP_OK;
}

void *mfp_buffer_ptr(mfp_buffer b) { return b ? b->data : nullptr; }
size_t mfp_buffer_size(mfp_buffer b) { return b ? b->size : 0; }

int mfp_buffer_destroy(mfp_buffer b)
{
  if (!b) return MFP_OK;
  dfree(b->data);
  delete b;
  return MFP_OK;
}

} // extern "C"
