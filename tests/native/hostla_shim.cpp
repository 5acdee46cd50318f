// C entry points over the host linear algebra of the block solvers
// (nemo3d_amd/csrc/nh_hostla.h), for the CPU unit tests.
#include "../../nemo3d_amd/csrc/nh_hostla.h"

using nh::BandT;
using nh::cd;

static BandT make(int p, int m, const double *A, const double *B) {
  BandT T;
  T.p = p;
  T.m = m;
  for (int k = 0; k < m; ++k) {
    std::vector<cd> a(p * p), b(p * p);
    for (int t = 0; t < p * p; ++t) {
      a[t] = cd(A[2 * (k * p * p + t)], A[2 * (k * p * p + t) + 1]);
      b[t] = cd(B[2 * (k * p * p + t)], B[2 * (k * p * p + t) + 1]);
    }
    T.A.push_back(a);
    T.B.push_back(b);
  }
  return T;
}

extern "C" {

int hl_band_count(int p, int m, const double *A, const double *B, double sigma) {
  std::vector<cd> work;
  return nh::band_count(make(p, m, A, B), sigma, work);
}

void hl_band_eigvec(int p, int m, const double *A, const double *B,
                    double lambda, double *x) {
  std::vector<cd> v;
  nh::band_eigvec(make(p, m, A, B), lambda, v);
  for (size_t i = 0; i < v.size(); ++i) {
    x[2 * i] = v[i].real();
    x[2 * i + 1] = v[i].imag();
  }
}

void hl_dense_eigh(int s, const double *H, double *w, double *V) {
  std::vector<cd> h(s * s);
  for (int t = 0; t < s * s; ++t) h[t] = cd(H[2 * t], H[2 * t + 1]);
  std::vector<double> ww;
  std::vector<std::vector<cd>> vv;
  nh::dense_eigh(s, h, ww, vv);
  for (int k = 0; k < s; ++k) {
    w[k] = ww[k];
    for (int i = 0; i < s; ++i) {
      V[2 * (k * s + i)] = vv[k][i].real();
      V[2 * (k * s + i) + 1] = vv[k][i].imag();
    }
  }
}

int hl_cholesky_upper(int p, double *G) {
  std::vector<cd> g(p * p);
  for (int t = 0; t < p * p; ++t) g[t] = cd(G[2 * t], G[2 * t + 1]);
  const bool ok = nh::cholesky_upper(p, g);
  for (int t = 0; t < p * p; ++t) {
    G[2 * t] = g[t].real();
    G[2 * t + 1] = g[t].imag();
  }
  return ok;
}
void hl_qr_r(int rows, int cols, const double *A, double *R) {
  std::vector<cd> a((size_t)rows * cols);
  for (size_t t = 0; t < a.size(); ++t) a[t] = cd(A[2 * t], A[2 * t + 1]);
  const std::vector<cd> r = nh::qr_r(rows, cols, a);
  for (size_t t = 0; t < r.size(); ++t) {
    R[2 * t] = r[t].real();
    R[2 * t + 1] = r[t].imag();
  }
}

void hl_refined_ritz(int m, int q, const double *T, const double *Rw,
                     double theta, double *y) {
  std::vector<cd> t((size_t)m * m), rw((size_t)q * q), yy(m);
  for (size_t k = 0; k < t.size(); ++k) t[k] = cd(T[2 * k], T[2 * k + 1]);
  for (size_t k = 0; k < rw.size(); ++k) rw[k] = cd(Rw[2 * k], Rw[2 * k + 1]);
  for (int k = 0; k < m; ++k) yy[k] = cd(y[2 * k], y[2 * k + 1]);
  nh::refined_ritz(m, q, t, rw, theta, yy);
  for (int k = 0; k < m; ++k) {
    y[2 * k] = yy[k].real();
    y[2 * k + 1] = yy[k].imag();
  }
}
}
