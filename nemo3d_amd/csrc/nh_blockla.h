// Block linear algebra shared by the block solvers: FP64 block Gram and
// block update kernels on the device, and small host routines (Cholesky,
// triangular inverse, Hermitian band matrices: Sturm/LDL^H inertia counts,
// band LU inverse iteration).  A dense Hermitian matrix is the band case
// with a single block.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "nh_hostla.h"
#include "nh_internal.h"

namespace nh {

constexpr int kMaxP = 16;

// ---------------- device block kernels ------------------------------------

// part[blk][a*P + b] (re, im) = sum_i conj(X_a[i]) Y_b[i] over this block,
// a < qa, b < qb (the other entries are zero)
template <int P>
__global__ __launch_bounds__(256) void k_gram(
    int64_t n, int qa, int qb, const double2 *__restrict__ X, int64_t ldx,
    const double2 *__restrict__ Y, int64_t ldy, double *__restrict__ part) {
  double re[P * P], im[P * P];
#pragma unroll
  for (int t = 0; t < P * P; ++t) re[t] = im[t] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 x[P], y[P];
#pragma unroll
    for (int a = 0; a < P; ++a) x[a] = a < qa ? X[a * ldx + i] : make_double2(0, 0);
#pragma unroll
    for (int b = 0; b < P; ++b) y[b] = b < qb ? Y[b * ldy + i] : make_double2(0, 0);
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int b = 0; b < P; ++b) {
        re[a * P + b] += x[a].x * y[b].x + x[a].y * y[b].y;
        im[a * P + b] += x[a].x * y[b].y - x[a].y * y[b].x;
      }
  }
  __shared__ double sh[4][2 * P * P];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int t = 0; t < P * P; ++t) {
    double a = re[t], b = im[t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off, 64);
      b += __shfl_down(b, off, 64);
    }
    if (lane == 0) {
      sh[w][2 * t] = a;
      sh[w][2 * t + 1] = b;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 2 * P * P; t += blockDim.x)
    part[(size_t)blockIdx.x * 2 * P * P + t] =
        sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
}

// out[t] = sum_k part[k * len + t] for t < len.  One workgroup per output
// (grid-stride over t), its threads split the k range; blockDim a multiple
// of 64, at most 256.  Launch with a grid of sum_cols_grid(len).
static __global__ void k_sum_cols(int nparts, int len, const double *__restrict__ part,
                                  double *__restrict__ out) {
  __shared__ double sh[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int t = blockIdx.x; t < len; t += gridDim.x) {
    double s = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x)
      s += part[(size_t)k * len + t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) sh[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = 0.0;
      for (int v = 0; v < (int)(blockDim.x >> 6); ++v) a += sh[v];
      out[t] = a;
    }
    __syncthreads();
  }
}
inline int sum_cols_grid(int len) { return std::min(std::max(len, 1), 4096); }

// Tiled block Gram over row chunks: tile t = (tx, ty) covers columns
// [tx P, tx P + P) of X and [ty P, ty P + P) of Y;
// part[(c * ntiles + t) * 2 P^2 + 2 (a P + b)] (re, im) = sum over the rows of
// chunk c of conj(X_a) Y_b.  Workgroup b computes tile b % ntiles of chunk
// b / ntiles, so the tiles of one chunk run side by side and the rows of Y
// they share come from the caches.
template <int P>
__global__ __launch_bounds__(256) void k_gram_tiles(
    int64_t n, int64_t chunk, int qx, int qy, const double2 *__restrict__ X,
    int64_t ldx, const double2 *__restrict__ Y, int64_t ldy,
    double *__restrict__ part) {
  const int ntx = (qx + P - 1) / P, nty = (qy + P - 1) / P, nt = ntx * nty;
  const int t = blockIdx.x % nt;
  const int64_t c = blockIdx.x / nt;
  const int tx = t / nty, ty = t % nty;
  const int qa = min(P, qx - tx * P), qb = min(P, qy - ty * P);
  const double2 *Xt = X + (int64_t)tx * P * ldx;
  const double2 *Yt = Y + (int64_t)ty * P * ldy;
  const int64_t i1 = min(n, (c + 1) * chunk);
  double re[P * P], im[P * P];
#pragma unroll
  for (int e = 0; e < P * P; ++e) re[e] = im[e] = 0.0;
  for (int64_t i = c * chunk + threadIdx.x; i < i1; i += blockDim.x) {
    double2 x[P], y[P];
#pragma unroll
    for (int a = 0; a < P; ++a) x[a] = a < qa ? Xt[a * ldx + i] : make_double2(0, 0);
#pragma unroll
    for (int b = 0; b < P; ++b) y[b] = b < qb ? Yt[b * ldy + i] : make_double2(0, 0);
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int b = 0; b < P; ++b) {
        re[a * P + b] += x[a].x * y[b].x + x[a].y * y[b].y;
        im[a * P + b] += x[a].x * y[b].y - x[a].y * y[b].x;
      }
  }
  __shared__ double sh[4][2 * P * P];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int e = 0; e < P * P; ++e) {
    double a = re[e], b = im[e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off, 64);
      b += __shfl_down(b, off, 64);
    }
    if (lane == 0) {
      sh[w][2 * e] = a;
      sh[w][2 * e + 1] = b;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * P * P; e += blockDim.x)
    part[(size_t)blockIdx.x * 2 * P * P + e] = sh[0][e] + sh[1][e] + sh[2][e] + sh[3][e];
}

// Y_c = beta Y_c + sum_a X_a M[a][c] for c < qy <= QY and a < qx <=
// kGemmMaxX in one pass over the rows (M complex, row-major qx x qy); Y must
// not overlap X.
constexpr int kGemmMaxX = 128;
template <int QY>
__global__ __launch_bounds__(256) void k_gemm_wide(
    int64_t n, int qx, int qy, const double2 *__restrict__ X, int64_t ldx,
    double2 *__restrict__ Y, int64_t ldy, const double2 *__restrict__ M,
    double beta) {
  __shared__ double2 ms[kGemmMaxX * QY];
  const int qx8 = (qx + 7) / 8 * 8;
  for (int t = threadIdx.x; t < qx8 * QY; t += blockDim.x) {
    const int a = t / QY, c = t % QY;
    ms[t] = a < qx && c < qy ? M[a * qy + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 acc[QY];
#pragma unroll
    for (int c = 0; c < QY; ++c) {
      acc[c] = make_double2(0.0, 0.0);
      if (beta != 0.0 && c < qy) {
        const double2 y = Y[c * ldy + i];
        acc[c] = make_double2(beta * y.x, beta * y.y);
      }
    }
    for (int a0 = 0; a0 < qx8; a0 += 8) {
      double2 x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        x[j] = a0 + j < qx ? X[(a0 + j) * ldx + i] : make_double2(0.0, 0.0);
#pragma unroll
      for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int c = 0; c < QY; ++c) {
          const double2 m = ms[(a0 + j) * QY + c];
          acc[c].x += x[j].x * m.x - x[j].y * m.y;
          acc[c].y += x[j].x * m.y + x[j].y * m.x;
        }
    }
#pragma unroll
    for (int c = 0; c < QY; ++c)
      if (c < qy) Y[c * ldy + i] = acc[c];
  }
}

// Y_c = beta * Y_c + sum_k X_k M[k][c]   (M complex, row-major qx x qy)
template <int P>
__global__ __launch_bounds__(256) void k_block_gemm(
    int64_t n, int qx, int qy, const double2 *__restrict__ X, int64_t ldx,
    double2 *__restrict__ Y, int64_t ldy, const double *__restrict__ M,
    double beta) {
  __shared__ double ms[2 * P * P];
  for (int t = threadIdx.x; t < 2 * qx * qy; t += blockDim.x) ms[t] = M[t];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 x[P];
#pragma unroll
    for (int k = 0; k < P; ++k) x[k] = k < qx ? X[k * ldx + i] : make_double2(0, 0);
#pragma unroll
    for (int c = 0; c < P; ++c) {
      if (c >= qy) break;
      double2 acc = beta == 0.0 ? make_double2(0, 0) : Y[c * ldy + i];
      acc.x *= beta;
      acc.y *= beta;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        if (k >= qx) break;
        const double mr = ms[2 * (k * qy + c)], mi = ms[2 * (k * qy + c) + 1];
        acc.x += x[k].x * mr - x[k].y * mi;
        acc.y += x[k].x * mi + x[k].y * mr;
      }
      Y[c * ldy + i] = acc;
    }
  }
}

}  // namespace nh

namespace nh {

// part[blk][2j..2j+1] = sum_i conj(X_j[i]) Y_j[i], j < q (column-wise dots)
template <int P>
__global__ __launch_bounds__(256) void k_coldot(
    int64_t n, int q, const double2 *__restrict__ X, int64_t ldx,
    const double2 *__restrict__ Y, int64_t ldy, double *__restrict__ part) {
  double re[P], im[P];
#pragma unroll
  for (int j = 0; j < P; ++j) re[j] = im[j] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      if (j < q) {
        const double2 x = X[j * ldx + i], y = Y[j * ldy + i];
        re[j] += x.x * y.x + x.y * y.y;
        im[j] += x.x * y.y - x.y * y.x;
      }
    }
  }
  __shared__ double sh[4][2 * P];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int j = 0; j < P; ++j) {
    double a = re[j], b = im[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_down(a, off, 64);
      b += __shfl_down(b, off, 64);
    }
    if (lane == 0) {
      sh[w][2 * j] = a;
      sh[w][2 * j + 1] = b;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 2 * P; t += blockDim.x)
    part[(size_t)blockIdx.x * 2 * P + t] = sh[0][t] + sh[1][t] + sh[2][t] + sh[3][t];
}

// Y_j = a_j X_j + b_j Y_j with real per-column scalars (ab[2j], ab[2j+1]);
// columns with a_j == 0 and b_j == 1 are left untouched.
template <int P>
__global__ __launch_bounds__(256) void k_col_axpby(
    int64_t n, int q, const double2 *__restrict__ X, int64_t ldx,
    double2 *__restrict__ Y, int64_t ldy, const double *__restrict__ ab) {
  __shared__ double s[2 * P];
  for (int t = threadIdx.x; t < 2 * q; t += blockDim.x) s[t] = ab[t];
  __syncthreads();
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      if (j >= q) break;
      const double a = s[2 * j], b = s[2 * j + 1];
      if (a == 0.0 && b == 1.0) continue;
      const double2 x = X[j * ldx + i];
      double2 y = Y[j * ldy + i];
      y.x = a * x.x + b * y.x;
      y.y = a * x.y + b * y.y;
      Y[j * ldy + i] = y;
    }
  }
}

}  // namespace nh
