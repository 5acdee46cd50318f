// Vector kernels of the solvers: dot products and norms with FP64
// accumulation (wavefront shuffles, then one partial per workgroup and a
// single-workgroup final sum, so results are deterministic), and axpby.
#include <mutex>

#include "nh_internal.h"
#include "nh_vector.h"

namespace nh {

// Scratch of nh_dot_noalloc: the partial sums of the dot product and its
// result, in a file-scope device array (each device holds its own copy
// until the process ends).  Host code uses the address that
// hipGetSymbolAddress gives for the current device, never the array's name.
constexpr int kNoAllocParts = kRedMaxBlocks;
static __device__ __align__(256) double scratch[2 * kNoAllocParts + 2];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch[kMaxDevices] = {};

// Resolves the device address of `scratch` on `device` (current device);
// called once per device by nh_context_create.
nh_status init_noalloc_scratch(int device) {
  NH_CHECK(device >= 0 && device < kMaxDevices, NH_ERR_ARG,
           "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch[device]) return NH_OK;
  void *ptr = nullptr;
  NH_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  NH_CHECK(ptr, NH_ERR_HIP, "hipGetSymbolAddress(scratch) gave no address");
  g_scratch[device] = static_cast<double *>(ptr);
  return NH_OK;
}

double *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}

nh_status ensure_reduce(nh_context *c, int blocks) {
  if (blocks <= c->red_cap) return NH_OK;
  (void)hipFree(c->red);
  (void)hipHostFree(c->hred);
  c->red = nullptr;
  c->hred = nullptr;
  const int cap = blocks < 4096 ? 4096 : blocks;
  NH_TRY(device_malloc(c, (void **)&c->red, (size_t)(cap + 64) * 2 * sizeof(double)));
  NH_HIP(hipHostMalloc(&c->hred, 64 * sizeof(double)));
  c->red_cap = cap;
  return NH_OK;
}

namespace {

__global__ __launch_bounds__(kRedThreads) void k_dot_partial(
    int64_t n, const double2 *__restrict__ x, const double2 *__restrict__ y,
    double *__restrict__ part) {
  double re = 0.0, im = 0.0;
  for (int64_t k = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kRedThreads) {
    const double2 a = x[k], b = y[k];
    re += a.x * b.x + a.y * b.y;
    im += a.x * b.y - a.y * b.x;
  }
  block_sum2(re, im);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = re;
    part[2 * blockIdx.x + 1] = im;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_nrm2_partial(
    int64_t n, const double2 *__restrict__ x, double *__restrict__ part) {
  double s = 0.0, z = 0.0;
  for (int64_t k = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kRedThreads) {
    const double2 a = x[k];
    s += a.x * a.x + a.y * a.y;
  }
  block_sum2(s, z);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s;
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

__global__ void k_axpby(int64_t n, double ar, double ai,
                        const double2 *__restrict__ x, double br, double bi,
                        double2 *__restrict__ y) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = x[k], b = y[k];
    y[k] = make_double2(ar * a.x - ai * a.y + br * b.x - bi * b.y,
                        ar * a.y + ai * a.x + br * b.y + bi * b.x);
  }
}

}  // namespace

int reduce_blocks(int64_t n) {
  int64_t b = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
  if (b > kRedMaxBlocks) b = kRedMaxBlocks;
  return (int)(b < 1 ? 1 : b);
}

nh_status reduce_dot(nh_context *c, int64_t n, const nh_cplx *x,
                     const nh_cplx *y, double *out2) {
  const int nb = reduce_blocks(n);
  NH_TRY(ensure_reduce(c, nb));
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kRedThreads), 0, c->stream,
                     n, reinterpret_cast<const double2 *>(x),
                     reinterpret_cast<const double2 *>(y), c->red);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, c->stream,
                     nb, c->red, c->red + 2 * (size_t)c->red_cap);
  NH_HIP(hipGetLastError());
  NH_HIP(hipMemcpyAsync(c->hred, c->red + 2 * (size_t)c->red_cap,
                        2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NH_HIP(hipStreamSynchronize(c->stream));
  out2[0] = c->hred[0];
  out2[1] = c->hred[1];
  return NH_OK;
}

// reduce_dot with the partial sums and the result in the device's file-scope
// scratch (no allocation); the result is copied to the host out2.
nh_status reduce_dot_noalloc(nh_context *c, int64_t n, const nh_cplx *x,
                             const nh_cplx *y, double *out2) {
  double *ws = noalloc_scratch(c->device);
  NH_CHECK(ws, NH_ERR_STATE, "nh_dot_noalloc: scratch of device %d not set up",
           c->device);
  const int nb = reduce_blocks(n);
  NH_CHECK(nb <= kNoAllocParts, NH_ERR_ARG,
           "nh_dot_noalloc: %d partial sums exceed the scratch (%d)", nb,
           kNoAllocParts);
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kRedThreads), 0, c->stream,
                     n, reinterpret_cast<const double2 *>(x),
                     reinterpret_cast<const double2 *>(y), ws);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, c->stream,
                     nb, ws, ws + 2 * (size_t)kNoAllocParts);
  NH_HIP(hipGetLastError());
  NH_HIP(hipMemcpyAsync(out2, ws + 2 * (size_t)kNoAllocParts, 2 * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  NH_HIP(hipStreamSynchronize(c->stream));
  return NH_OK;
}

nh_status reduce_nrm2sq(nh_context *c, int64_t n, const nh_cplx *x,
                        double *out) {
  const int nb = reduce_blocks(n);
  NH_TRY(ensure_reduce(c, nb));
  hipLaunchKernelGGL(k_nrm2_partial, dim3(nb), dim3(kRedThreads), 0, c->stream,
                     n, reinterpret_cast<const double2 *>(x), c->red);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0, c->stream,
                     nb, c->red, c->red + 2 * (size_t)c->red_cap);
  NH_HIP(hipGetLastError());
  NH_HIP(hipMemcpyAsync(c->hred, c->red + 2 * (size_t)c->red_cap,
                        2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  NH_HIP(hipStreamSynchronize(c->stream));
  *out = c->hred[0];
  return NH_OK;
}

nh_status launch_axpby(nh_context *c, int64_t n, nh_cplx a, const nh_cplx *x,
                       nh_cplx b, nh_cplx *y) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_axpby, dim3((int)(g < 1 ? 1 : g)), dim3(256), 0,
                     c->stream, n, a.re, a.im,
                     reinterpret_cast<const double2 *>(x), b.re, b.im,
                     reinterpret_cast<double2 *>(y));
  NH_HIP(hipGetLastError());
  return NH_OK;
}

}  // namespace nh
