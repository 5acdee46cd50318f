// Shared reduction helpers (each translation unit gets its own copy of the
// kernels: the library is built without relocatable device code).
#pragma once

#include <hip/hip_runtime.h>

namespace nh {

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 2048;

int reduce_blocks(int64_t n);

// Sum of two FP64 values over a workgroup of kRedThreads; thread 0 holds
// the result.  Order of additions is fixed, so results are reproducible.
__device__ __forceinline__ void block_sum2(double &a, double &b) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  __shared__ double sa[kRedThreads / 64], sb[kRedThreads / 64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tb = 0.0;
#pragma unroll
    for (int k = 0; k < kRedThreads / 64; ++k) {
      ta += sa[k];
      tb += sb[k];
    }
    a = ta;
    b = tb;
  }
}

// block_sum2 over kRedThreads threads with caller-provided LDS scratch
// (2 * kRedThreads / 64 doubles, free for reuse: a barrier comes first).
__device__ __forceinline__ void block_sum2_lds(double &a, double &b,
                                               double *scratch) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  __syncthreads();
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    scratch[w] = a;
    scratch[kRedThreads / 64 + w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tb = 0.0;
#pragma unroll
    for (int k = 0; k < kRedThreads / 64; ++k) {
      ta += scratch[k];
      tb += scratch[kRedThreads / 64 + k];
    }
    a = ta;
    b = tb;
  }
}

// Final sum of `nparts` (re, im) partials into out[0..1].
static __global__ __launch_bounds__(kRedThreads) void k_sum_partials(
    int nparts, const double *__restrict__ part, double *__restrict__ out) {
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nparts; k += kRedThreads) {
    a += part[2 * k];
    b += part[2 * k + 1];
  }
  block_sum2(a, b);
  if (threadIdx.x == 0) {
    out[0] = a;
    out[1] = b;
  }
}

}  // namespace nh
