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

#include "nh_internal.h"

namespace nh {

using cd = std::complex<double>;
constexpr int kMaxP = 16;

// ---------------- device block kernels ------------------------------------

// part[blk][a*q + b] (re, im) = sum_i conj(X_a[i]) Y_b[i] over this block
template <int P>
__global__ __launch_bounds__(256) void k_gram(
    int64_t n, int q, const double2 *__restrict__ X, int64_t ldx,
    const double2 *__restrict__ Y, int64_t ldy, double *__restrict__ part) {
  double re[P * P], im[P * P];
#pragma unroll
  for (int t = 0; t < P * P; ++t) re[t] = im[t] = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 x[P], y[P];
#pragma unroll
    for (int a = 0; a < P; ++a) x[a] = a < q ? X[a * ldx + i] : make_double2(0, 0);
#pragma unroll
    for (int b = 0; b < P; ++b) y[b] = b < q ? Y[b * ldy + i] : make_double2(0, 0);
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

static __global__ void k_sum_cols(int nparts, int len, const double *__restrict__ part,
                           double *__restrict__ out) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < len;
       t += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < nparts; ++k) s += part[(size_t)k * len + t];
    out[t] = s;
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

// ---------------- host: small dense and band linear algebra ----------------

// Cholesky G = R^H R (R upper); returns false if not positive definite.
inline bool cholesky_upper(int p, std::vector<cd> &G) {
  for (int j = 0; j < p; ++j) {
    double d = G[j * p + j].real();
    for (int k = 0; k < j; ++k) d -= std::norm(G[k * p + j]);
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    G[j * p + j] = d;
    for (int c = j + 1; c < p; ++c) {
      cd s = G[j * p + c];
      for (int k = 0; k < j; ++k) s -= std::conj(G[k * p + j]) * G[k * p + c];
      G[j * p + c] = s / d;
    }
    for (int r = j + 1; r < p; ++r) G[r * p + j] = 0.0;
  }
  return true;
}

// inverse of upper triangular R
inline std::vector<cd> inv_upper(int p, const std::vector<cd> &R) {
  std::vector<cd> X(p * p, 0.0);
  for (int c = 0; c < p; ++c) {
    X[c * p + c] = 1.0 / R[c * p + c];
    for (int r = c - 1; r >= 0; --r) {
      cd s = 0.0;
      for (int k = r + 1; k <= c; ++k) s += R[r * p + k] * X[k * p + c];
      X[r * p + c] = -s / R[r * p + r];
    }
  }
  return X;
}

inline std::vector<cd> matmul(int p, const std::vector<cd> &A, const std::vector<cd> &B) {
  std::vector<cd> C(p * p, 0.0);
  for (int i = 0; i < p; ++i)
    for (int k = 0; k < p; ++k)
      for (int j = 0; j < p; ++j) C[i * p + j] += A[i * p + k] * B[k * p + j];
  return C;
}

// Hermitian band matrix from blocks: element (r, c), c >= r, |c - r| <= p.
struct BandT {
  int p = 0, m = 0;  // block width, number of blocks
  std::vector<std::vector<cd>> A, B;  // A_j (Hermitian), B_j (upper)
  int n() const { return m * p; }
  cd at(int r, int c) const {  // r <= c
    const int br = r / p, bc = c / p;
    if (bc == br) return A[br][(r % p) * p + (c % p)];
    if (bc == br + 1) return std::conj(B[br][(c % p) * p + (r % p)]);  // B^H
    return 0.0;
  }
};

// number of eigenvalues of T (first m blocks) below sigma: LDL^H inertia
inline int band_count(const BandT &T, double sigma, std::vector<cd> &work) {
  const int n = T.n(), p = T.p, w = p + 1;
  work.assign((size_t)n * w, 0.0);
  for (int r = 0; r < n; ++r)
    for (int d = 0; d <= p && r + d < n; ++d)
      work[(size_t)r * w + d] = T.at(r, r + d) - (d == 0 ? sigma : 0.0);
  int neg = 0;
  for (int k = 0; k < n; ++k) {
    double dk = work[(size_t)k * w].real();
    if (std::fabs(dk) < 1e-300) dk = 1e-300;
    if (dk < 0) ++neg;
    for (int i = 1; i <= p && k + i < n; ++i) {
      const cd aki = work[(size_t)k * w + i];
      if (aki == 0.0) continue;
      const cd f = std::conj(aki) / dk;
      for (int j = i; j <= p && k + j < n; ++j)
        work[(size_t)(k + i) * w + (j - i)] -= f * work[(size_t)k * w + j];
    }
  }
  return neg;
}

// (T - lambda) x = b by band LU with partial pivoting; x overwrites b.
inline void band_solve(const BandT &T, double lambda, std::vector<cd> &b) {
  const int n = T.n(), p = T.p;
  const int kl = p, ku = p, W = 2 * kl + ku + 1;  // LAPACK-style band
  std::vector<cd> ab((size_t)n * W, 0.0);
  auto AB = [&](int i, int j) -> cd & { return ab[(size_t)j * W + (kl + ku + i - j)]; };
  for (int r = 0; r < n; ++r)
    for (int c = std::max(0, r - p); c <= std::min(n - 1, r + p); ++c) {
      const cd v = c >= r ? T.at(r, c) : std::conj(T.at(c, r));
      AB(r, c) = v - (r == c ? lambda + 1e-13 * (std::fabs(lambda) + 1.0) : 0.0);
    }
  std::vector<int> piv(n);
  for (int j = 0; j < n; ++j) {
    const int km = std::min(kl, n - 1 - j);
    int ip = j;
    double best = std::abs(AB(j, j));
    for (int i = j + 1; i <= j + km; ++i)
      if (std::abs(AB(i, j)) > best) best = std::abs(AB(i, j)), ip = i;
    piv[j] = ip;
    const int ju = std::min(n - 1, j + kl + ku);
    if (ip != j)
      for (int c = j; c <= ju; ++c) std::swap(AB(j, c), AB(ip, c));
    if (std::abs(AB(j, j)) < 1e-300) AB(j, j) = 1e-300;
    for (int i = j + 1; i <= j + km; ++i) {
      const cd f = AB(i, j) / AB(j, j);
      AB(i, j) = f;
      for (int c = j + 1; c <= ju; ++c) AB(i, c) -= f * AB(j, c);
    }
  }
  for (int j = 0; j < n; ++j) {  // L y = P b
    if (piv[j] != j) std::swap(b[j], b[piv[j]]);
    const int km = std::min(kl, n - 1 - j);
    for (int i = j + 1; i <= j + km; ++i) b[i] -= AB(i, j) * b[j];
  }
  for (int j = n - 1; j >= 0; --j) {  // U x = y
    cd s = b[j];
    for (int c = j + 1; c <= std::min(n - 1, j + kl + ku); ++c) s -= AB(j, c) * b[c];
    b[j] = s / AB(j, j);
  }
}

// Inverse iteration; `against` holds already accepted vectors of the same
// (degenerate) eigenvalue, which the result is kept orthogonal to.
inline void band_eigvec(const BandT &T, double lambda, std::vector<cd> &x,
                 uint64_t seed = 0,
                 const std::vector<const std::vector<cd> *> &against = {}) {
  const int n = T.n();
  x.assign(n, 0.0);
  uint64_t s = 0x2545F4914F6CDD1Dull ^ (seed * 0x9E3779B97F4A7C15ull);
  for (int i = 0; i < n; ++i) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    const double a = (double)(s >> 11) / 9007199254740992.0;
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    x[i] = cd(0.5 + a, (double)(s >> 11) / 9007199254740992.0 - 0.5);
  }
  for (int it = 0; it < 3; ++it) {
    band_solve(T, lambda, x);
    for (int pass = 0; pass < 2; ++pass)
      for (const auto *v : against) {
        cd d = 0.0;
        for (int i = 0; i < n; ++i) d += std::conj((*v)[i]) * x[i];
        for (int i = 0; i < n; ++i) x[i] -= d * (*v)[i];
      }
    double nrm = 0.0;
    for (auto &v : x) nrm += std::norm(v);
    nrm = std::sqrt(nrm);
    for (auto &v : x) v /= nrm;
  }
}


// Eigenpairs of a small dense Hermitian matrix (row-major s x s) with
// eigenvalues in ascending order: Sturm bisection on the one-block band
// form, then inverse iteration (degenerate clusters orthogonalised).
inline void dense_eigh(int s, const std::vector<cd> &Hm, std::vector<double> &w,
                       std::vector<std::vector<cd>> &V) {
  BandT T;
  T.p = s;
  T.m = 1;
  T.A.push_back(Hm);
  double g = 0.0;
  for (int r = 0; r < s; ++r) {
    double t = 0.0;
    for (int c = 0; c < s; ++c) t += std::abs(Hm[r * s + c]);
    g = std::max(g, t);
  }
  std::vector<cd> work;
  w.assign(s, 0.0);
  V.assign(s, {});
  for (int k = 0; k < s; ++k) {
    double a = -g - 1.0, b = g + 1.0;
    for (int it = 0; it < 200 && b - a > 1e-15 * (g + 1.0); ++it) {
      const double mid = 0.5 * (a + b);
      if (band_count(T, mid, work) > k) b = mid;
      else a = mid;
    }
    w[k] = 0.5 * (a + b);
  }
  for (int k = 0; k < s; ++k) {
    std::vector<const std::vector<cd> *> cluster;
    for (int f = 0; f < k; ++f)
      if (std::fabs(w[f] - w[k]) <= 1e-10 * (g + 1.0)) cluster.push_back(&V[f]);
    band_eigvec(T, w[k], V[k], (uint64_t)k + 7, cluster);
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
