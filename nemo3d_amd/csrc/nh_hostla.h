// Small host linear algebra of the block solvers: Cholesky, triangular
// inverse, Hermitian band matrices (LDL^H inertia counts for Sturm
// bisection, band LU inverse iteration) and dense Hermitian eigenpairs (the
// one-block band case).  Plain C++ (no HIP), so it is unit-tested on the CPU.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

namespace nh {

using cd = std::complex<double>;

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

// First column j of a q-column block, given its Gram matrix G = W^H W,
// whose norm after orthogonalisation against columns 0..j-1 (the Cholesky
// diagonal) is below tol, or below 1e-6 of its own norm (the Gram matrix
// resolves no smaller fraction: its entries carry rounding of order
// 1e-16 |w|^2); -1 if none.
inline int cholesky_first_small(int q, std::vector<cd> G, double tol) {
  for (int j = 0; j < q; ++j) {
    double d = G[j * q + j].real();
    const double g = d;
    for (int k = 0; k < j; ++k) d -= std::norm(G[k * q + j]);
    if (!(d > tol * tol) || !(d > 1e-12 * g)) return j;
    d = std::sqrt(d);
    G[j * q + j] = d;
    for (int c = j + 1; c < q; ++c) {
      cd s = G[j * q + c];
      for (int k = 0; k < j; ++k) s -= std::conj(G[k * q + j]) * G[k * q + c];
      G[j * q + c] = s / d;
    }
  }
  return -1;
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
// Hermitian block tridiagonal T of m blocks: A_j (s_j x s_j Hermitian,
// row-major) on the diagonal, B_j (s_{j+1} x s_j, row-major) below it.
// Blocks are all p wide unless `sz` holds the sizes (a block Lanczos run
// whose block shrank by deflation); call layout() after changing m or sz.
// The nonzeros of every row lie within p of the diagonal (p = the largest
// block; B_j is upper triangular up to the rows a deflation removed).
struct BandT {
  int p = 0, m = 0;  // bandwidth (largest block), number of blocks
  std::vector<std::vector<cd>> A, B;
  std::vector<int> sz;  // block sizes (empty: every block p)
  std::vector<int> off_, blk_;
  void layout() {
    off_.assign(m + 1, 0);
    for (int j = 0; j < m; ++j) off_[j + 1] = off_[j] + size(j);
    blk_.assign(off_[m], 0);
    for (int j = 0; j < m; ++j)
      for (int r = off_[j]; r < off_[j + 1]; ++r) blk_[r] = j;
  }
  int size(int j) const { return sz.empty() ? p : sz[j]; }
  int offset(int j) const { return sz.empty() ? j * p : off_[j]; }
  int n() const { return sz.empty() ? m * p : off_[m]; }
  cd at(int r, int c) const {  // r <= c
    const int br = sz.empty() ? r / p : blk_[r];
    const int bc = sz.empty() ? c / p : blk_[c];
    const int lr = r - offset(br), lc = c - offset(bc), sr = size(br);
    if (bc == br) return A[br][lr * sr + lc];
    if (bc == br + 1) return std::conj(B[br][lc * sr + lr]);  // B^H
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

// R factor (cols x cols, row-major upper) of the Householder QR of A
// (rows x cols, row-major, rows >= cols).
inline std::vector<cd> qr_r(int rows, int cols, std::vector<cd> A) {
  for (int j = 0; j < cols; ++j) {
    double nrm = 0.0;
    for (int r = j; r < rows; ++r) nrm += std::norm(A[r * cols + j]);
    nrm = std::sqrt(nrm);
    if (nrm == 0.0) continue;
    const cd a = A[j * cols + j];
    const cd ph = std::abs(a) > 0.0 ? a / std::abs(a) : cd(1.0);
    const cd alpha = -ph * nrm;  // R_jj; v = x - alpha e1
    std::vector<cd> v(rows - j);
    for (int r = j; r < rows; ++r) v[r - j] = A[r * cols + j];
    v[0] -= alpha;
    double vn = 0.0;
    for (const auto &z : v) vn += std::norm(z);
    if (vn == 0.0) continue;
    for (int c = j; c < cols; ++c) {
      cd d = 0.0;
      for (int r = j; r < rows; ++r) d += std::conj(v[r - j]) * A[r * cols + c];
      d *= 2.0 / vn;
      for (int r = j; r < rows; ++r) A[r * cols + c] -= d * v[r - j];
    }
  }
  std::vector<cd> R((size_t)cols * cols, 0.0);
  for (int r = 0; r < cols; ++r)
    for (int c = r; c < cols; ++c) R[(size_t)r * cols + c] = A[(size_t)r * cols + c];
  return R;
}

// Refined Ritz vector of a Krylov basis B (m columns, orthonormal) with
// H B = B T + W E^T, W orthogonal to B, W^H W = Rw^H Rw (Rw q x q upper)
// and E^T selecting the last q coefficients: the unit y minimising
// ||(H - theta) B y||^2 = ||(T - theta) y||^2 + ||Rw y_last||^2, by
// inverse iteration on the R factor of [T - theta I; Rw E^T] starting from
// y (in/out).  T: m x m row-major (the Arnoldi matrix, not symmetrised).
inline void refined_ritz(int m, int q, const std::vector<cd> &T,
                         const std::vector<cd> &Rw, double theta,
                         std::vector<cd> &y, int iters = 3) {
  const int rows = m + q;
  std::vector<cd> A((size_t)rows * m, 0.0);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c)
      A[(size_t)r * m + c] = T[(size_t)r * m + c] - (r == c ? theta : 0.0);
  for (int r = 0; r < q; ++r)
    for (int c = 0; c < q; ++c) A[(size_t)(m + r) * m + m - q + c] = Rw[r * q + c];
  std::vector<cd> R = qr_r(rows, m, A);
  double dmax = 0.0;
  for (int i = 0; i < m; ++i) dmax = std::max(dmax, std::abs(R[(size_t)i * m + i]));
  const double floor_ = 1e-15 * std::max(dmax, 1e-300);
  for (int i = 0; i < m; ++i)
    if (std::abs(R[(size_t)i * m + i]) < floor_) R[(size_t)i * m + i] = floor_;
  for (int it = 0; it < iters; ++it) {
    // R^H z = y (lower), then R y' = z (upper)
    std::vector<cd> z(m);
    for (int i = 0; i < m; ++i) {
      cd s = y[i];
      for (int k = 0; k < i; ++k) s -= std::conj(R[(size_t)k * m + i]) * z[k];
      z[i] = s / std::conj(R[(size_t)i * m + i]);
    }
    for (int i = m - 1; i >= 0; --i) {
      cd s = z[i];
      for (int k = i + 1; k < m; ++k) s -= R[(size_t)i * m + k] * y[k];
      y[i] = s / R[(size_t)i * m + i];
    }
    double n = 0.0;
    for (const auto &v : y) n += std::norm(v);
    n = std::sqrt(n);
    for (auto &v : y) v /= n;
  }
}

}  // namespace nh
