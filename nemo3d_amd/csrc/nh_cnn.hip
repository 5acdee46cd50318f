// 3-D convolutional network of the learned warm start: forward, backward,
// Adam and the eigenvector-subspace loss, all on the device.
//
// Network: L conv layers (3x3x3, zero padding, ReLU) c_in -> W -> ... -> W,
// then a 1x1x1 head W -> c_out.  Activations are FP32, channels last
// ([point][channel], point = (x*gy + y)*gz + z).  Weights are kept as FP32
// master copies for Adam and used as BF16 in the forward and backward
// products (FP32 accumulation).
// Parameter vector: for each conv layer W_l[co][ci][27], b_l[co]; then the
// head W_h[co][ci], b_h[co].
//
// Loss of one training structure with target eigenvectors v_f (FP64, from
// the plain solver) and predicted vectors y_e (the head output mapped to
// atoms): L = mean_e (1 - ||P_e y_e||^2 / ||y_e||^2), P_e the projector on
// the degenerate cluster of target e (a Kramers pair in the spin model).
// It is invariant to the phase of each state and to rotations inside a
// degenerate pair.
#include <chrono>
#include <cstring>
#include <vector>

#include "nh_blockla.h"
#include "nh_vector.h"

struct nh_cnn {
  nh_context *ctx = nullptr;
  int device = 0;  // the context may be destroyed before the network
  int c_in = 0, width = 0, layers = 0, c_out = 0;
  int64_t n_params = 0;
  std::vector<int64_t> w_off, b_off;  // per layer (conv layers, then head)
  float *master = nullptr;            // FP32 parameters
  uint16_t *bf = nullptr;             // BF16 copy used by the products
  float *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
  int64_t step = 0;
  // activations for the current grid
  int64_t cap_points = 0;
  std::vector<float *> pre, act;  // per conv layer, [P][W]
  float *in = nullptr, *out = nullptr, *dact = nullptr, *dtmp = nullptr;
  double *part = nullptr;
};

namespace nh {
namespace {

__device__ __forceinline__ float bf16_to_f(uint16_t h) {
  return __uint_as_float(((uint32_t)h) << 16);
}
// round to nearest even
__device__ __forceinline__ uint16_t f_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (uint16_t)(u >> 16);
}

__global__ void k_to_bf16(int64_t n, const float *__restrict__ a,
                          uint16_t *__restrict__ b) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    b[k] = f_to_bf16(a[k]);
}

// neighbour offset k in 0..26 -> (dx, dy, dz) in {-1,0,1}
__device__ __forceinline__ void koff(int k, int &dx, int &dy, int &dz) {
  dx = k / 9 - 1;
  dy = (k / 3) % 3 - 1;
  dz = k % 3 - 1;
}

// forward conv3x3x3 + ReLU: one thread per point, all CO outputs
template <int CO>
__global__ __launch_bounds__(128) void k_conv_fwd(
    int gx, int gy, int gz, int ci_n, const float *__restrict__ in,
    const uint16_t *__restrict__ W, const uint16_t *__restrict__ b,
    float *__restrict__ pre, float *__restrict__ act) {
  extern __shared__ uint16_t wsh[];  // [CO][ci_n][27]
  const int nw = CO * ci_n * 27;
  for (int t = threadIdx.x; t < nw; t += blockDim.x) wsh[t] = W[t];
  __syncthreads();
  const int64_t P = (int64_t)gx * gy * gz;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int z = p % gz, y = (p / gz) % gy, x = p / ((int64_t)gy * gz);
  float acc[CO];
#pragma unroll
  for (int co = 0; co < CO; ++co) acc[co] = bf16_to_f(b[co]);
  for (int k = 0; k < 27; ++k) {
    int dx, dy, dz;
    koff(k, dx, dy, dz);
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    if (xx < 0 || yy < 0 || zz < 0 || xx >= gx || yy >= gy || zz >= gz) continue;
    const float *src = in + (((int64_t)xx * gy + yy) * gz + zz) * ci_n;
    for (int ci = 0; ci < ci_n; ++ci) {
      const float v = src[ci];
      if (v == 0.0f) continue;
#pragma unroll
      for (int co = 0; co < CO; ++co)
        acc[co] += bf16_to_f(wsh[(co * ci_n + ci) * 27 + k]) * v;
    }
  }
#pragma unroll
  for (int co = 0; co < CO; ++co) {
    pre[p * CO + co] = acc[co];
    act[p * CO + co] = acc[co] > 0.0f ? acc[co] : 0.0f;
  }
}

// head 1x1x1: out[p][co] = b[co] + sum_ci W[co][ci] a[p][ci].  Lane =
// output channel of a 64-channel chunk (blockIdx.y), so a wavefront writes
// 64 consecutive outputs of one point and reads that point's inputs as one
// broadcast; each wavefront walks groups of kHeadPts points.  The chunk's weights sit
// in LDS as FP32, [ci][64] (lane-consecutive, no bank conflicts).
constexpr int kHeadPts = 8;
__global__ __launch_bounds__(256) void k_head_fwd(
    int64_t P, int ci_n, int co_n, const float *__restrict__ a,
    const uint16_t *__restrict__ W, const uint16_t *__restrict__ b,
    float *__restrict__ out) {
  extern __shared__ float whs[];  // [ci_n][64]
  const int co0 = blockIdx.y * 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int t = threadIdx.x; t < 64 * ci_n; t += blockDim.x) {
    const int ci = t >> 6, c = t & 63;
    whs[t] = co0 + c < co_n ? bf16_to_f(W[(size_t)(co0 + c) * ci_n + ci]) : 0.0f;
  }
  __syncthreads();
  const int co = co0 + lane;
  const bool ok = co < co_n;
  const float bias = ok ? bf16_to_f(b[co]) : 0.0f;
  // groups of kHeadPts points, grid-stride (the grid is capped)
  for (int64_t g = (int64_t)blockIdx.x * 4 + w; g * kHeadPts < P;
       g += (int64_t)gridDim.x * 4) {
    for (int k = 0; k < kHeadPts; ++k) {
      const int64_t p = g * kHeadPts + k;
      if (p >= P) break;
      const float *ap = a + p * ci_n;
      float s = bias;
      for (int ci = 0; ci < ci_n; ++ci) s += whs[ci * 64 + lane] * ap[ci];
      if (ok) out[p * co_n + co] = s;
    }
  }
}

// head backward, data: da[p][ci] = sum_co W[co][ci] dout[p][co].  Lane =
// input channel (ci_n <= 64 lanes of a wavefront per point), so the
// point's dout is read as broadcasts and da written contiguously; W is
// read lane-consecutive from the caches; grid-stride over points.
__global__ __launch_bounds__(256) void k_head_bwd_data(
    int64_t P, int ci_n, int co_n, const float *__restrict__ dout,
    const uint16_t *__restrict__ W, float *__restrict__ da) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane >= ci_n) return;
  for (int64_t p = (int64_t)blockIdx.x * 4 + w; p < P; p += (int64_t)gridDim.x * 4) {
    const float *dp = dout + p * co_n;
    float s = 0.0f;
    for (int co = 0; co < co_n; ++co) s += bf16_to_f(W[(size_t)co * ci_n + lane]) * dp[co];
    da[p * ci_n + lane] = s;
  }
}

// sum over a workgroup of 256 (float accumulate in double for stability)
__device__ __forceinline__ double wg_sum256(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __shared__ double s[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) s[w] = v;
  __syncthreads();
  return s[0] + s[1] + s[2] + s[3];
}

// head backward, weights: one workgroup per output channel co
__global__ __launch_bounds__(256) void k_head_bwd_w(
    int64_t P, int ci_n, int co_n, const float *__restrict__ dout,
    const float *__restrict__ a, float *__restrict__ gW, float *__restrict__ gb) {
  const int co = blockIdx.x;
  double acc[64];
  for (int ci = 0; ci < ci_n; ++ci) acc[ci] = 0.0;
  double sb = 0.0;
  for (int64_t p = threadIdx.x; p < P; p += blockDim.x) {
    const float g = dout[p * co_n + co];
    if (g == 0.0f) continue;
    sb += g;
    for (int ci = 0; ci < ci_n; ++ci) acc[ci] += (double)g * a[p * ci_n + ci];
  }
  for (int ci = 0; ci < ci_n; ++ci) {
    const double t = wg_sum256(acc[ci]);
    if (threadIdx.x == 0) gW[(size_t)co * ci_n + ci] = (float)t;
  }
  const double t = wg_sum256(sb);
  if (threadIdx.x == 0) gb[co] = (float)t;
}

// dpre = dact * relu'(pre)
__global__ void k_relu_bwd(int64_t n, const float *__restrict__ pre,
                           const float *__restrict__ dact, float *__restrict__ dpre) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    dpre[k] = pre[k] > 0.0f ? dact[k] : 0.0f;
}

// conv backward, data: din[p][ci] = sum_{co,k} W[co][ci][k] dpre[p - d_k][co]
template <int CO>
__global__ __launch_bounds__(128) void k_conv_bwd_data(
    int gx, int gy, int gz, int ci_n, const float *__restrict__ dpre,
    const uint16_t *__restrict__ W, float *__restrict__ din) {
  extern __shared__ uint16_t wsh[];
  const int nw = CO * ci_n * 27;
  for (int t = threadIdx.x; t < nw; t += blockDim.x) wsh[t] = W[t];
  __syncthreads();
  const int64_t P = (int64_t)gx * gy * gz;
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int z = p % gz, y = (p / gz) % gy, x = p / ((int64_t)gy * gz);
  for (int ci = 0; ci < ci_n; ++ci) {
    float s = 0.0f;
    for (int k = 0; k < 27; ++k) {
      int dx, dy, dz;
      koff(k, dx, dy, dz);
      const int xx = x - dx, yy = y - dy, zz = z - dz;
      if (xx < 0 || yy < 0 || zz < 0 || xx >= gx || yy >= gy || zz >= gz) continue;
      const float *g = dpre + (((int64_t)xx * gy + yy) * gz + zz) * CO;
#pragma unroll
      for (int co = 0; co < CO; ++co)
        s += bf16_to_f(wsh[(co * ci_n + ci) * 27 + k]) * g[co];
    }
    din[p * ci_n + ci] = s;
  }
}

// conv backward, weights: one workgroup per (co, ci); 27 sums over points
__global__ __launch_bounds__(256) void k_conv_bwd_w(
    int gx, int gy, int gz, int ci_n, int co_n, const float *__restrict__ dpre,
    const float *__restrict__ in, float *__restrict__ gW) {
  const int co = blockIdx.x / ci_n, ci = blockIdx.x % ci_n;
  const int64_t P = (int64_t)gx * gy * gz;
  double acc[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = 0.0;
  for (int64_t p = threadIdx.x; p < P; p += blockDim.x) {
    const float g = dpre[p * co_n + co];
    if (g == 0.0f) continue;
    const int z = p % gz, y = (p / gz) % gy, x = p / ((int64_t)gy * gz);
#pragma unroll
    for (int k = 0; k < 27; ++k) {
      const int dx = k / 9 - 1, dy = (k / 3) % 3 - 1, dz = k % 3 - 1;
      const int xx = x + dx, yy = y + dy, zz = z + dz;
      if (xx < 0 || yy < 0 || zz < 0 || xx >= gx || yy >= gy || zz >= gz) continue;
      acc[k] += (double)g * in[(((int64_t)xx * gy + yy) * gz + zz) * ci_n + ci];
    }
  }
  for (int k = 0; k < 27; ++k) {
    const double t = wg_sum256(acc[k]);
    if (threadIdx.x == 0) gW[((size_t)co * ci_n + ci) * 27 + k] = (float)t;
  }
}

// bias gradient of a conv layer: one workgroup per channel
__global__ __launch_bounds__(256) void k_bias_bwd(int64_t P, int co_n,
                                                  const float *__restrict__ dpre,
                                                  float *__restrict__ gb) {
  const int co = blockIdx.x;
  double s = 0.0;
  for (int64_t p = threadIdx.x; p < P; p += blockDim.x) s += dpre[p * co_n + co];
  const double t = wg_sum256(s);
  if (threadIdx.x == 0) gb[co] = (float)t;
}

__global__ void k_adam(int64_t n, float *__restrict__ w,
                       const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
  }
}

// head output -> start block: y_e[atom*nb + o] from out[point][(e*2nb + slot*nb + o)*2 .. +1]
__global__ void k_map_out(int64_t A, int nb, int ne, int c_out,
                          const int32_t *__restrict__ point,
                          const int32_t *__restrict__ slot,
                          const float *__restrict__ out, double2 *__restrict__ Y,
                          int64_t ld) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < A * nb;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = t / nb;
    const int o = t % nb;
    const float *src = out + (int64_t)point[a] * c_out;
    for (int e = 0; e < ne; ++e) {
      const int ch = 2 * ((e * 2 + slot[a]) * nb + o);
      Y[e * ld + t] = make_double2(src[ch], src[ch + 1]);
    }
  }
}

// gradient w.r.t. the start block -> head output gradient (points without
// atoms get zero: the caller clears dout first)
__global__ void k_map_grad(int64_t A, int nb, int ne, int c_out,
                           const int32_t *__restrict__ point,
                           const int32_t *__restrict__ slot,
                           const double2 *__restrict__ G, int64_t ld,
                           float *__restrict__ dout) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < A * nb;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t a = t / nb;
    const int o = t % nb;
    float *dst = dout + (int64_t)point[a] * c_out;
    for (int e = 0; e < ne; ++e) {
      const int ch = 2 * ((e * 2 + slot[a]) * nb + o);
      const double2 g = G[e * ld + t];
      dst[ch] = (float)g.x;
      dst[ch + 1] = (float)g.y;
    }
  }
}

int g1(int64_t n, int b = 256) {
  int64_t g = (n + b - 1) / b;
  return (int)(g < 1 ? 1 : (g > 65535 ? 65535 : g));
}

nh_status ensure_acts(nh_cnn *c, int64_t P) {
  if (P <= c->cap_points) return NH_OK;
  for (float *q : c->pre) (void)hipFree(q);
  for (float *q : c->act) (void)hipFree(q);
  (void)hipFree(c->in);
  (void)hipFree(c->out);
  (void)hipFree(c->dact);
  (void)hipFree(c->dtmp);
  c->pre.assign(c->layers, nullptr);
  c->act.assign(c->layers, nullptr);
  for (int l = 0; l < c->layers; ++l) {
    NH_TRY(device_malloc(c->ctx, (void **)&c->pre[l], P * c->width * sizeof(float)));
    NH_TRY(device_malloc(c->ctx, (void **)&c->act[l], P * c->width * sizeof(float)));
  }
  const int cmax = std::max(std::max(c->c_in, c->width), c->c_out);
  NH_TRY(device_malloc(c->ctx, (void **)&c->in, P * c->c_in * sizeof(float)));
  NH_TRY(device_malloc(c->ctx, (void **)&c->out, P * c->c_out * sizeof(float)));
  NH_TRY(device_malloc(c->ctx, (void **)&c->dact, P * cmax * sizeof(float)));
  NH_TRY(device_malloc(c->ctx, (void **)&c->dtmp, P * cmax * sizeof(float)));
  c->cap_points = P;
  return NH_OK;
}

nh_status forward(nh_cnn *c, int gx, int gy, int gz) {
  const int64_t P = (int64_t)gx * gy * gz;
  hipStream_t st = c->ctx->stream;
  const float *src = c->in;
  int ci = c->c_in;
  for (int l = 0; l < c->layers; ++l) {
    const size_t shm = (size_t)c->width * ci * 27 * sizeof(uint16_t);
    NH_CHECK(c->width == 32, NH_ERR_ARG, "network width must be 32");
    hipLaunchKernelGGL(k_conv_fwd<32>, dim3(g1(P, 128)), dim3(128), shm, st, gx,
                       gy, gz, ci, src, c->bf + c->w_off[l], c->bf + c->b_off[l],
                       c->pre[l], c->act[l]);
    src = c->act[l];
    ci = c->width;
  }
  const int L = c->layers;
  hipLaunchKernelGGL(k_head_fwd, dim3(g1(P, 4 * kHeadPts), (c->c_out + 63) / 64),
                     dim3(256), 64 * ci * sizeof(float), st, P, ci, c->c_out,
                     src, c->bf + c->w_off[L], c->bf + c->b_off[L], c->out);
  NH_HIP(hipGetLastError());
  return NH_OK;
}

// parameter gradient of sum(out * dout); dout in c->dtmp layout [P][c_out]
nh_status backward(nh_cnn *c, int gx, int gy, int gz, const float *dout) {
  const int64_t P = (int64_t)gx * gy * gz;
  hipStream_t st = c->ctx->stream;
  const int L = c->layers, W = c->width;
  const float *alast = L > 0 ? c->act[L - 1] : c->in;
  hipLaunchKernelGGL(k_head_bwd_w, dim3(c->c_out), dim3(256), 0, st, P, W,
                     c->c_out, dout, alast, c->grad + c->w_off[L],
                     c->grad + c->b_off[L]);
  NH_CHECK(W <= 64, NH_ERR_ARG, "network width above 64");
  hipLaunchKernelGGL(k_head_bwd_data, dim3(g1(P, 4)), dim3(256), 0, st, P, W,
                     c->c_out, dout, c->bf + c->w_off[L], c->dact);
  for (int l = L - 1; l >= 0; --l) {
    const int ci = l == 0 ? c->c_in : W;
    const float *in = l == 0 ? c->in : c->act[l - 1];
    // dact (for act[l]) -> dpre in dtmp
    hipLaunchKernelGGL(k_relu_bwd, dim3(g1(P * W)), dim3(256), 0, st, P * W,
                       c->pre[l], c->dact, c->dtmp);
    hipLaunchKernelGGL(k_conv_bwd_w, dim3(W * ci), dim3(256), 0, st, gx, gy, gz,
                       ci, W, c->dtmp, in, c->grad + c->w_off[l]);
    hipLaunchKernelGGL(k_bias_bwd, dim3(W), dim3(256), 0, st, P, W, c->dtmp,
                       c->grad + c->b_off[l]);
    if (l > 0) {
      const size_t shm = (size_t)W * ci * 27 * sizeof(uint16_t);
      hipLaunchKernelGGL(k_conv_bwd_data<32>, dim3(g1(P, 128)), dim3(128), shm,
                         st, gx, gy, gz, ci, c->dtmp, c->bf + c->w_off[l], c->dact);
    }
  }
  NH_HIP(hipGetLastError());
  return NH_OK;
}

void init_layout(nh_cnn *c) {
  c->w_off.clear();
  c->b_off.clear();
  int64_t off = 0;
  int ci = c->c_in;
  for (int l = 0; l < c->layers; ++l) {
    c->w_off.push_back(off);
    off += (int64_t)c->width * ci * 27;
    c->b_off.push_back(off);
    off += c->width;
    ci = c->width;
  }
  c->w_off.push_back(off);
  off += (int64_t)c->c_out * ci;
  c->b_off.push_back(off);
  off += c->c_out;
  c->n_params = off;
}

nh_status sync_bf16(nh_cnn *c) {
  hipLaunchKernelGGL(k_to_bf16, dim3(g1(c->n_params)), dim3(256), 0,
                     c->ctx->stream, c->n_params, c->master, c->bf);
  NH_HIP(hipGetLastError());
  return NH_OK;
}

nh_status upload_feats(nh_cnn *c, const float *feat, int64_t P) {
  NH_TRY(ensure_acts(c, P));
  NH_HIP(hipMemcpyAsync(c->in, feat, P * c->c_in * sizeof(float),
                        hipMemcpyHostToDevice, c->ctx->stream));
  return NH_OK;
}

}  // namespace
}  // namespace nh

using namespace nh;

extern "C" {

nh_status nh_cnn_create(nh_context *ctx, int c_in, int width, int layers,
                        int c_out, uint64_t seed, nh_cnn **out) {
  NH_CHECK(ctx && out && c_in > 0 && c_in <= 64 && width == 32 && layers >= 1 &&
               layers <= 8 && c_out > 0,
           NH_ERR_ARG, "nh_cnn_create: bad shape");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  nh_cnn *c = new nh_cnn();
  c->ctx = ctx;
  c->device = ctx->device;
  c->c_in = c_in;
  c->width = width;
  c->layers = layers;
  c->c_out = c_out;
  init_layout(c);
  std::vector<float> w(c->n_params, 0.0f);
  uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
  auto gauss = [&]() {
    double u[2];
    for (double &v : u) {
      s = s * 6364136223846793005ull + 1442695040888963407ull;
      v = ((double)(s >> 11) + 0.5) / 9007199254740992.0;
    }
    return std::sqrt(-2.0 * std::log(u[0])) * std::cos(6.283185307179586 * u[1]);
  };
  int ci = c_in;
  for (int l = 0; l <= layers; ++l) {
    const bool head = l == layers;
    const int fan = head ? ci : ci * 27;
    const int co = head ? c_out : width;
    const double sd = std::sqrt(2.0 / fan) * (head ? 0.5 : 1.0);
    for (int64_t k = 0; k < (int64_t)co * fan; ++k) w[c->w_off[l] + k] = (float)(sd * gauss());
    ci = width;
  }
  nh_status st = NH_OK;
  auto mk = [&](void **p, size_t bytes) {
    if (st == NH_OK && device_malloc(c->ctx, p, bytes) != NH_OK) {
      set_error("nh_cnn_create: out of device memory");
      st = NH_ERR_NOMEM;
    }
  };
  mk((void **)&c->master, c->n_params * sizeof(float));
  mk((void **)&c->bf, c->n_params * sizeof(uint16_t));
  mk((void **)&c->grad, c->n_params * sizeof(float));
  mk((void **)&c->m1, c->n_params * sizeof(float));
  mk((void **)&c->m2, c->n_params * sizeof(float));
  if (st == NH_OK) {
    hipError_t e = hipMemcpy(c->master, w.data(), c->n_params * sizeof(float),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(c->m1, 0, c->n_params * sizeof(float));
    if (e == hipSuccess) e = hipMemset(c->m2, 0, c->n_params * sizeof(float));
    if (e == hipSuccess) e = hipMemset(c->grad, 0, c->n_params * sizeof(float));
    if (e != hipSuccess) {
      nh_cnn_destroy(c);
      NH_CHECK(false, NH_ERR_HIP, "network set-up: %s", hipGetErrorString(e));
    }
    st = sync_bf16(c);
  }
  if (st != NH_OK) {
    nh_cnn_destroy(c);
    return st;
  }
  *out = c;
  return NH_OK;
}

void nh_cnn_destroy(nh_cnn *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  for (float *q : c->pre) (void)hipFree(q);
  for (float *q : c->act) (void)hipFree(q);
  void *ps[] = {c->master, c->bf, c->grad, c->m1, c->m2, c->in, c->out,
                c->dact, c->dtmp, c->part};
  for (void *q : ps)
    if (q) (void)hipFree(q);
  delete c;
}

int64_t nh_cnn_num_params(const nh_cnn *c) { return c ? c->n_params : 0; }

nh_status nh_cnn_get_params(nh_cnn *c, float *host) {
  NH_CHECK(c && host, NH_ERR_ARG, "nh_cnn_get_params: bad arguments");
  NH_HIP(hipMemcpy(host, c->master, c->n_params * sizeof(float), hipMemcpyDeviceToHost));
  return NH_OK;
}

nh_status nh_cnn_set_params(nh_cnn *c, const float *host) {
  NH_CHECK(c && host, NH_ERR_ARG, "nh_cnn_set_params: bad arguments");
  NH_HIP(hipMemcpy(c->master, host, c->n_params * sizeof(float), hipMemcpyHostToDevice));
  NH_TRY(sync_bf16(c));
  NH_HIP(hipStreamSynchronize(c->ctx->stream));
  return NH_OK;
}

nh_status nh_cnn_forward(nh_cnn *c, const float *feat, int gx, int gy, int gz,
                         float *out) {
  NH_CHECK(c && feat && out && gx > 0 && gy > 0 && gz > 0, NH_ERR_ARG,
           "nh_cnn_forward: bad arguments");
  (void)hipGetLastError();
  const int64_t P = (int64_t)gx * gy * gz;
  NH_TRY(upload_feats(c, feat, P));
  NH_TRY(forward(c, gx, gy, gz));
  NH_HIP(hipMemcpyAsync(out, c->out, P * c->c_out * sizeof(float),
                        hipMemcpyDeviceToHost, c->ctx->stream));
  NH_HIP(hipStreamSynchronize(c->ctx->stream));
  return NH_OK;
}

nh_status nh_cnn_backward(nh_cnn *c, const float *feat, int gx, int gy, int gz,
                          const float *dout, float *grad) {
  NH_CHECK(c && feat && dout && grad, NH_ERR_ARG, "nh_cnn_backward: bad arguments");
  (void)hipGetLastError();
  const int64_t P = (int64_t)gx * gy * gz;
  NH_TRY(upload_feats(c, feat, P));
  NH_TRY(forward(c, gx, gy, gz));
  float *d = c->out;  // reuse the output buffer for dout
  NH_HIP(hipMemcpyAsync(d, dout, P * c->c_out * sizeof(float),
                        hipMemcpyHostToDevice, c->ctx->stream));
  NH_TRY(backward(c, gx, gy, gz, d));
  NH_HIP(hipMemcpyAsync(grad, c->grad, c->n_params * sizeof(float),
                        hipMemcpyDeviceToHost, c->ctx->stream));
  NH_HIP(hipStreamSynchronize(c->ctx->stream));
  return NH_OK;
}

nh_status nh_cnn_bind(nh_cnn *c, nh_context *ctx) {
  NH_CHECK(c && ctx, NH_ERR_ARG, "nh_cnn_bind: null argument");
  NH_CHECK(ctx->device == c->device, NH_ERR_ARG,
           "nh_cnn_bind: context on another device");
  // every network call synchronises its stream before returning, so no
  // work is pending on the previous context's stream
  c->ctx = ctx;
  return NH_OK;
}

nh_status nh_cnn_start_block(nh_cnn *c, const float *feat, int gx, int gy,
                             int gz, const int32_t *atom_point,
                             const int32_t *atom_slot, int n_states,
                             nh_cplx *block, int ptr_kind) {
  NH_CHECK(c && feat && atom_point && atom_slot && block && n_states > 0 &&
               n_states * 2 * c->ctx->nb * 2 <= c->c_out,
           NH_ERR_ARG, "nh_cnn_start_block: bad arguments");
  (void)hipGetLastError();
  nh_context *x = c->ctx;
  const int64_t P = (int64_t)gx * gy * gz, A = x->n_atoms;
  NH_TRY(upload_feats(c, feat, P));
  NH_TRY(forward(c, gx, gy, gz));
  int32_t *dp = nullptr, *ds = nullptr;
  NH_TRY(device_malloc(x, (void **)&dp, A * sizeof(int32_t)));
  NH_TRY(device_malloc(x, (void **)&ds, A * sizeof(int32_t)));
  {
    hipError_t e = hipMemcpy(dp, atom_point, A * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(ds, atom_slot, A * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(dp);
      (void)hipFree(ds);
      NH_CHECK(false, NH_ERR_HIP, "nh_cnn_start_block: %s", hipGetErrorString(e));
    }
  }
  double2 *Y = reinterpret_cast<double2 *>(block);
  double2 *tmp = nullptr;
  if (ptr_kind == NH_PTR_HOST) {
    if (device_malloc(x, (void **)&tmp, (size_t)n_states * x->n * sizeof(double2)) != NH_OK) {
      (void)hipFree(dp);
      (void)hipFree(ds);
      NH_CHECK(false, NH_ERR_NOMEM, "nh_cnn_start_block: out of memory");
    }
    Y = tmp;
  }
  hipLaunchKernelGGL(k_map_out, dim3(g1(A * x->nb)), dim3(256), 0, x->stream, A,
                     x->nb, n_states, c->c_out, dp, ds, c->out, Y, x->n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && tmp)
    e = hipMemcpyAsync(block, tmp, (size_t)n_states * x->n * sizeof(double2),
                       hipMemcpyDeviceToHost, x->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(x->stream);
  (void)hipFree(dp);
  (void)hipFree(ds);
  if (tmp) (void)hipFree(tmp);
  NH_CHECK(e == hipSuccess, NH_ERR_HIP, "nh_cnn_start_block: %s", hipGetErrorString(e));
  return NH_OK;
}

nh_status nh_cnn_train_step(nh_cnn *c, const float *feat, int gx, int gy, int gz,
                            const int32_t *atom_point, const int32_t *atom_slot,
                            const nh_cplx *targets, int n_targets,
                            const int32_t *cluster, int n_states,
                            const int32_t *out_cluster, double lr, double *loss,
                            float *grad) {
  NH_CHECK(c && feat && atom_point && atom_slot && targets && cluster &&
               out_cluster && loss && n_targets >= 1 && n_targets <= 16 &&
               n_states >= 1 && n_states <= 8 &&
               n_states * 2 * c->ctx->nb * 2 <= c->c_out,
           NH_ERR_ARG, "nh_cnn_train_step: bad arguments");
  (void)hipGetLastError();
  nh_context *x = c->ctx;
  const int64_t P = (int64_t)gx * gy * gz, A = x->n_atoms, N = x->n;
  const int ne = n_states;
  NH_TRY(upload_feats(c, feat, P));
  NH_TRY(forward(c, gx, gy, gz));
  // device scratch: point/slot maps, Y (prediction, 8 columns), V (targets,
  // 16 columns), G (gradient, 8 columns)
  std::vector<void *> mem;
  auto mk = [&](void **p, size_t bytes) -> nh_status {
    NH_TRY(device_malloc(x, p, bytes));
    mem.push_back(*p);
    return NH_OK;
  };
  auto cleanup = [&]() {
    for (void *q : mem) (void)hipFree(q);
  };
  int32_t *dp, *ds;
  double2 *Y, *V, *G;
  double *part, *sum, *dM;
  nh_status st = NH_OK;
  if (st == NH_OK) st = mk((void **)&dp, A * sizeof(int32_t));
  if (st == NH_OK) st = mk((void **)&ds, A * sizeof(int32_t));
  if (st == NH_OK) st = mk((void **)&Y, (size_t)8 * N * sizeof(double2));
  if (st == NH_OK) st = mk((void **)&V, (size_t)16 * N * sizeof(double2));
  if (st == NH_OK) st = mk((void **)&G, (size_t)8 * N * sizeof(double2));
  if (st == NH_OK) st = mk((void **)&part, (size_t)512 * 2 * 64 * sizeof(double));
  if (st == NH_OK) st = mk((void **)&sum, 2 * 64 * sizeof(double));
  if (st == NH_OK) st = mk((void **)&dM, 2 * 64 * sizeof(double));
  if (st != NH_OK) {
    cleanup();
    return st;
  }
  hipStream_t s = x->stream;
  // an asynchronous call that fails ends the step with the buffers released
#define TS_HIP(call)                                                            \
  do {                                                                          \
    const hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) {                                                     \
      cleanup();                                                                \
      NH_CHECK(false, NH_ERR_HIP, "train step: %s", hipGetErrorString(e_));    \
    }                                                                           \
  } while (0)
  TS_HIP(hipMemcpyAsync(dp, atom_point, A * sizeof(int32_t), hipMemcpyHostToDevice, s));
  TS_HIP(hipMemcpyAsync(ds, atom_slot, A * sizeof(int32_t), hipMemcpyHostToDevice, s));
  TS_HIP(hipMemsetAsync(Y, 0, (size_t)8 * N * sizeof(double2), s));
  TS_HIP(hipMemsetAsync(V, 0, (size_t)16 * N * sizeof(double2), s));
  // targets may be host or device memory (unified addressing)
  TS_HIP(hipMemcpyAsync(V, targets, (size_t)n_targets * N * sizeof(double2),
                 hipMemcpyDefault, s));
  hipLaunchKernelGGL(k_map_out, dim3(g1(A * x->nb)), dim3(256), 0, s, A, x->nb,
                     ne, c->c_out, dp, ds, c->out, Y, N);
  // C = V^H Y (16 x 8, two 8x8 halves) and the Gram Y^H Y
  std::vector<double> h(2 * 128), hb(2 * 64);
  for (int half = 0; half < 2; ++half) {
    hipLaunchKernelGGL((k_gram<8>), dim3(512), dim3(256), 0, s, N, 8, 8,
                       V + (int64_t)half * 8 * N, N, Y, N, part);
    hipLaunchKernelGGL(k_sum_cols, dim3(128), dim3(128), 0, s, 512, 128, part, sum);
    TS_HIP(hipMemcpyAsync(h.data() + half * 128, sum, 128 * sizeof(double),
                   hipMemcpyDeviceToHost, s));
    // the next k_sum_cols overwrites `sum`: wait for the copy
    TS_HIP(hipStreamSynchronize(s));
  }
  hipLaunchKernelGGL((k_gram<8>), dim3(512), dim3(256), 0, s, N, 8, 8, Y, N, Y, N, part);
  hipLaunchKernelGGL(k_sum_cols, dim3(128), dim3(128), 0, s, 512, 128, part, sum);
  TS_HIP(hipMemcpyAsync(hb.data(), sum, 128 * sizeof(double), hipMemcpyDeviceToHost, s));
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    cleanup();
    NH_CHECK(false, NH_ERR_HIP, "train step: %s", hipGetErrorString(e));
  }
  // C[f][e] = <v_f, y_e>
  auto C = [&](int f, int ee) {
    const double *q = h.data() + (f / 8) * 128 + 2 * ((f % 8) * 8 + ee);
    return cd(q[0], q[1]);
  };
  // L = mean_e (1 - a_e / b_e), a_e = sum over the targets of e's cluster of
  // |C[f][e]|^2, b_e = ||y_e||^2;  G_e = dL/dRe(y_e) + i dL/dIm(y_e)
  //     = 2 (-sum_f v_f C[f][e] / b_e + a_e y_e / b_e^2) / ne
  std::vector<cd> M(128, 0.0), D(64, 0.0);
  double L = 0.0;
  for (int ee = 0; ee < ne; ++ee) {
    const double b = std::max(hb[2 * (ee * 8 + ee)], 1e-300);
    double a = 0.0;
    for (int f = 0; f < n_targets; ++f)
      if (cluster[f] == out_cluster[ee]) a += std::norm(C(f, ee));
    L += (1.0 - a / b) / ne;
    for (int f = 0; f < n_targets; ++f)
      if (cluster[f] == out_cluster[ee]) M[f * 8 + ee] = -2.0 * C(f, ee) / (b * ne);
    D[ee * 8 + ee] = 2.0 * a / (b * b * ne);
  }
  *loss = L;
  std::vector<double> hm(128);
  for (int half = 0; half < 2; ++half) {
    for (int k = 0; k < 64; ++k) {
      hm[2 * k] = M[half * 64 + k].real();
      hm[2 * k + 1] = M[half * 64 + k].imag();
    }
    TS_HIP(hipMemcpyAsync(dM, hm.data(), 128 * sizeof(double), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL((k_block_gemm<8>), dim3(g1(N)), dim3(256), 0, s, N, 8, 8,
                       V + (int64_t)half * 8 * N, N, G, N, dM, half ? 1.0 : 0.0);
    TS_HIP(hipStreamSynchronize(s));  // hm is rewritten next
  }
  for (int k = 0; k < 64; ++k) {
    hm[2 * k] = D[k].real();
    hm[2 * k + 1] = 0.0;
  }
  TS_HIP(hipMemcpyAsync(dM, hm.data(), 128 * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((k_block_gemm<8>), dim3(g1(N)), dim3(256), 0, s, N, 8, 8, Y,
                     N, G, N, dM, 1.0);
  // dL/d(out): real and imaginary parts of the complex gradient
  TS_HIP(hipMemsetAsync(c->dtmp, 0, P * c->c_out * sizeof(float), s));
  hipLaunchKernelGGL(k_map_grad, dim3(g1(A * x->nb)), dim3(256), 0, s, A, x->nb,
                     ne, c->c_out, dp, ds, G, N, c->dtmp);
  // backward needs a dout buffer that survives: move it into c->out
  TS_HIP(hipMemcpyAsync(c->out, c->dtmp, P * c->c_out * sizeof(float),
                 hipMemcpyDeviceToDevice, s));
#undef TS_HIP
  e = hipStreamSynchronize(s);
  if (e == hipSuccess) {
    nh_status bs = backward(c, gx, gy, gz, c->out);
    if (bs != NH_OK) {
      cleanup();
      return bs;
    }
    if (grad)
      e = hipMemcpyAsync(grad, c->grad, c->n_params * sizeof(float),
                         hipMemcpyDeviceToHost, s);
  }
  if (e == hipSuccess && lr > 0.0) {
    c->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)c->step);
    const float c2 = 1.0f - std::pow(b2, (float)c->step);
    hipLaunchKernelGGL(k_adam, dim3(g1(c->n_params)), dim3(256), 0, s,
                       c->n_params, c->master, c->grad, c->m1, c->m2, (float)lr,
                       b1, b2, c1, c2, 1e-8f);
    sync_bf16(c);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  NH_CHECK(e == hipSuccess, NH_ERR_HIP, "train step: %s", hipGetErrorString(e));
  return NH_OK;
}

}  // extern "C"
