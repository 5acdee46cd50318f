// FP64 refinement of approximate eigenpairs, the last stage of the mixed
// precision solver: the Ritz vectors of a Lanczos run with FP32 products are
// eigenvectors of the FP32-rounded operator, so their FP64 residuals
// ||H x - e x|| sit at the level of the FP32 rounding of H x (broadband in
// the spectrum).  A few rounds of
//   V = [X, H X, H^2 X, ...]   (block Krylov, Arnoldi with classical
//                               Gram-Schmidt applied twice, Cholesky QR)
//   H V = V T + W E^T          (T from the Arnoldi coefficients)
//   x_j = V y_j, y_j minimising ||(T - e_j) y||^2 + ||Rw y_last||^2
//                              (refined Ritz vector, W^H W = Rw^H Rw)
// in FP64 reduce that error; e_j is updated to the Rayleigh quotient.  The
// basis lives in HBM ((depth + 1) * k vectors).
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "nh_blockla.h"
#include "nh_blockops.h"
#include "nh_internal.h"
#include "nh_vector.h"

namespace nh {
namespace {


class Refine {
 public:
  Refine(nh_context *c, int k, int depth, double tol, double hnorm)
      : c_(c), k_(k), d_(depth), kp_(k),  // blocks unpadded: the Gram and
                                             // update kernels take any width
        tol_(tol), hnorm_(hnorm), ops_(c, kp_, (depth + 1) * kp_) {
    n_ = c->n;
  }
  ~Refine() {
    for (void *q : mem_) pool_release(c_, q);
  }
  nh_status setup();
  nh_status round(std::vector<double> &target, std::vector<double> &resid,
                  int &matvecs);
  double2 *X() { return X_; }

 private:
  // from the context's workspace pool, zero-filled
  nh_status alloc(void **q, size_t bytes) {
    NH_TRY(pool_alloc(c_, q, bytes));
    mem_.push_back(*q);
    NH_HIP(hipMemsetAsync(*q, 0, bytes, c_->stream));
    return NH_OK;
  }
  double2 *blk(int i) { return V_ + (size_t)i * kp_ * n_; }
  // G (qx x qy, row-major) = X^H Y
  nh_status gram(const double2 *X, int qx, const double2 *Y, int qy,
                 std::vector<cd> &G) {
    return ops_.gram(X, qx, Y, qy, G);
  }
  // Y[:, 0..qy) = beta Y + X M, M (qx x qy, row-major)
  nh_status gemm(const double2 *X, int qx, double2 *Y, int qy,
                 const std::vector<cd> &M, double beta) {
    return ops_.gemm(X, qx, Y, qy, M, beta);
  }
  // Cholesky QR twice of the k columns of W into Q; R = R2 R1 (k x k)
  nh_status cholqr(double2 *W, double2 *Q, std::vector<cd> &R, bool &ok) {
    return ops_.cholqr(W, Q, k_, R, ok);
  }

  nh_context *c_;
  int k_, d_, kp_;
  double tol_, hnorm_;
  BlockOps ops_;

 public:
  int basis_blocks_ = 0;  // blocks used, summed over rounds

 private:
  int64_t n_ = 0;
  std::vector<void *> mem_;
  double2 *V_ = nullptr, *W_ = nullptr, *X_ = nullptr;
  double *part_ = nullptr, *sum_ = nullptr, *dM_ = nullptr;
  int nparts_ = 0;
};

nh_status Refine::setup() {
  NH_TRY(ops_.setup());
  nparts_ = reduce_blocks(n_);
  NH_TRY(alloc((void **)&V_, (size_t)(d_ + 1) * kp_ * n_ * sizeof(double2)));
  NH_TRY(alloc((void **)&W_, (size_t)kp_ * n_ * sizeof(double2)));
  NH_TRY(alloc((void **)&X_, (size_t)kp_ * n_ * sizeof(double2)));
  NH_TRY(alloc((void **)&part_, (size_t)nparts_ * 32 * sizeof(double)));
  NH_TRY(alloc((void **)&sum_, 32 * sizeof(double)));
  NH_TRY(alloc((void **)&dM_, 32 * sizeof(double)));
  return NH_OK;
}

nh_status Refine::round(std::vector<double> &target, std::vector<double> &resid,
                        int &matvecs) {
  const int k = k_, kp = kp_;
  const int64_t n = n_;
  // padded basis index b * kp + c (c < k valid); T over padded indices
  const int mp = (d_ + 1) * kp;
  std::vector<cd> T((size_t)mp * mp, 0.0);
  bool ok = true;
  std::vector<cd> R, R0;
  NH_TRY(cholqr(X_, blk(0), R0, ok));  // X = V_0 R0
  NH_CHECK(ok, NH_ERR_ARG, "refine: start vectors are linearly dependent");
  // Arnoldi matrix and Gram factor of the remainder for a basis of nb
  // blocks; refined Ritz vectors (y_j), values and the residual norms
  // ||(H - e) V y|| they give (exact up to rounding)
  std::vector<cd> Rw;
  auto small = [&](int nb, int outers, std::vector<cd> &M,
                   std::vector<double> &theta, std::vector<double> &sigma) {
    std::vector<int> idx;
    for (int b = 0; b < nb; ++b)
      for (int cc = 0; cc < k; ++cc) idx.push_back(b * kp + cc);
    const int m = (int)idx.size();
    std::vector<cd> Ta((size_t)m * m), Th((size_t)m * m);
    for (int r = 0; r < m; ++r)
      for (int cc = 0; cc < m; ++cc) {
        Ta[(size_t)r * m + cc] = T[(size_t)idx[r] * mp + idx[cc]];
        Th[(size_t)r * m + cc] = 0.5 * (T[(size_t)idx[r] * mp + idx[cc]] +
                                        std::conj(T[(size_t)idx[cc] * mp + idx[r]]));
      }
    M.assign((size_t)nb * kp * k, 0.0);
    theta.assign(k, 0.0);
    sigma.assign(k, 0.0);
    for (int j = 0; j < k; ++j) {
      std::vector<cd> y(m, 0.0);
      for (int a = 0; a < k; ++a) y[a] = R0[a * k + j];
      double th = target[j];
      for (int outer = 0; outer < outers; ++outer) {
        refined_ritz(m, k, Ta, Rw, th, y);
        cd rq = 0.0;
        for (int r = 0; r < m; ++r) {
          cd t = 0.0;
          for (int cc = 0; cc < m; ++cc) t += Th[(size_t)r * m + cc] * y[cc];
          rq += std::conj(y[r]) * t;
        }
        th = rq.real();
      }
      // ||(T - th) y||^2 + ||Rw y_last||^2
      double s2 = 0.0;
      for (int r = 0; r < m; ++r) {
        cd t = -th * y[r];
        for (int cc = 0; cc < m; ++cc) t += Ta[(size_t)r * m + cc] * y[cc];
        s2 += std::norm(t);
      }
      for (int r = 0; r < k; ++r) {
        cd t = 0.0;
        for (int cc = r; cc < k; ++cc) t += Rw[r * k + cc] * y[m - k + cc];
        s2 += std::norm(t);
      }
      theta[j] = th;
      sigma[j] = std::sqrt(s2);
      for (int r = 0; r < m; ++r) M[(size_t)idx[r] * k + j] = y[r];
    }
  };
  // the basis grows one block at a time; from 4 blocks on (every second
  // step) the refined residuals decide whether the next block is needed
  int nb = 1;  // blocks in the basis
  std::vector<cd> M;
  std::vector<double> theta, sigma;
  for (int i = 0; i <= d_; ++i) {
    NH_TRY(launch_hv(c_, 0.0, reinterpret_cast<const nh_cplx *>(blk(i)),
                     reinterpret_cast<nh_cplx *>(W_), k, n));
    matvecs += k;
    const int rows = (i + 1) * kp;
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<cd> C, Cm((size_t)rows * k);
      NH_TRY(gram(V_, rows, W_, kp, C));
      for (int r = 0; r < rows; ++r)
        for (int cc = 0; cc < k; ++cc) {
          T[(size_t)r * mp + i * kp + cc] += C[(size_t)r * kp + cc];
          Cm[(size_t)r * k + cc] = -C[(size_t)r * kp + cc];
        }
      NH_TRY(gemm(V_, rows, W_, k, Cm, 1.0));
    }
    // H V = V T + W E^T with W (remainder of block i) orthogonal to V: its
    // Gram factor Rw enters the residual norms of combinations of V
    std::vector<cd> G;
    NH_TRY(gram(W_, kp, W_, kp, G));
    Rw.assign((size_t)k * k, 0.0);
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b) Rw[a * k + b] = G[(size_t)a * kp + b];
    double gtr = 0.0;
    for (int a = 0; a < k; ++a) gtr += Rw[a * k + a].real();
    for (int a = 0; a < k; ++a) Rw[a * k + a] += 1e-30 * gtr + 1e-300;
    if (!cholesky_upper(k, Rw)) std::fill(Rw.begin(), Rw.end(), cd(0.0));
    if (i == d_) break;
    if (i >= 3 && (i & 1) == 1) {
      small(nb, 1, M, theta, sigma);
      bool enough = true;
      for (int j = 0; j < k; ++j)
        if (!(sigma[j] <= 0.8 * tol_ * hnorm_)) enough = false;
      if (enough) break;
    }
    NH_TRY(cholqr(W_, blk(i + 1), R, ok));
    if (!ok) break;  // Krylov space exhausted: W holds the tiny remainder
    for (int a = 0; a < k; ++a)
      for (int b = 0; b < k; ++b)
        T[(size_t)((i + 1) * kp + a) * mp + i * kp + b] = R[a * k + b];
    ++nb;
  }
  basis_blocks_ += nb;
  // refined Ritz vector per eigenvalue: the unit combination of the basis
  // with the smallest ||(H - e) x|| (interior eigenvalues: plain Ritz
  // vectors of a growing basis pick up spurious mixtures), e updated to its
  // Rayleigh quotient, twice; started from the current vector (V_0 R0 e_j)
  small(nb, 2, M, theta, sigma);
  for (int j = 0; j < k; ++j) target[j] = theta[j];
  // X = V Y (padded rows zero)
  NH_TRY(gemm(V_, nb * kp, X_, k, M, 0.0));
  // residuals ||H x - e x|| (explicit difference, no cancellation)
  NH_TRY(launch_hv(c_, 0.0, reinterpret_cast<const nh_cplx *>(X_),
                   reinterpret_cast<nh_cplx *>(W_), k, n));
  matvecs += k;
  std::vector<double> ab(2 * k);
  for (int j = 0; j < k; ++j) {
    ab[2 * j] = -target[j];
    ab[2 * j + 1] = 1.0;
  }
  NH_HIP(hipMemcpyAsync(dM_, ab.data(), ab.size() * sizeof(double),
                        hipMemcpyHostToDevice, c_->stream));
  int64_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  hipLaunchKernelGGL((k_col_axpby<16>), dim3((int)g), dim3(256), 0, c_->stream,
                     n, k, X_, n, W_, n, dM_);
  hipLaunchKernelGGL((k_coldot<16>), dim3(nparts_), dim3(256), 0, c_->stream,
                     n, k, W_, n, W_, n, part_);
  hipLaunchKernelGGL(k_sum_cols, dim3(32), dim3(256), 0, c_->stream, nparts_,
                     32, part_, sum_);
  NH_HIP(hipGetLastError());
  std::vector<double> h(32);
  NH_HIP(hipMemcpyAsync(h.data(), sum_, h.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c_->stream));
  NH_HIP(hipStreamSynchronize(c_->stream));
  resid.assign(k, 0.0);
  for (int j = 0; j < k; ++j) resid[j] = std::sqrt(std::max(h[2 * j], 0.0));
  return NH_OK;
}

}  // namespace
}  // namespace nh

using namespace nh;

extern "C" nh_status nh_refine_eigenpairs(nh_context *ctx,
                                          const nh_refine_opts *o, int32_t k,
                                          nh_cplx *X, int ptr_kind,
                                          double *eigenvalues,
                                          double *residuals, int32_t *rounds,
                                          int32_t *matvecs) {
  NH_CHECK(ctx && o && X && eigenvalues && residuals, NH_ERR_ARG,
           "refine: null argument");
  NH_CHECK(k >= 1 && k <= 16 && o->depth >= 1 && o->depth <= 32 &&
               o->max_rounds >= 1 && o->tolerance > 0 && o->hnorm > 0,
           NH_ERR_ARG, "refine: bad k %d / depth %d / rounds %d / tolerance",
           k, o->depth, o->max_rounds);
  NH_CHECK(ptr_kind == NH_PTR_HOST || ptr_kind == NH_PTR_DEVICE, NH_ERR_ARG,
           "refine: bad ptr_kind %d", ptr_kind);
  NH_CHECK(ctx->layout == NH_LAYOUT_BLOCKS || sk_like(ctx),
           NH_ERR_STATE, "refine: build the Hamiltonian first");
  NH_CHECK(ctx->precision == NH_PREC_FP64, NH_ERR_STATE,
           "refine: the operator must be FP64 (nh_ham_set_precision)");
  NH_HIP(hipSetDevice(ctx->device));
  Refine rf(ctx, k, o->depth, o->tolerance, o->hnorm);
  NH_TRY(rf.setup());
  const size_t bytes = (size_t)k * ctx->n * sizeof(double2);
  NH_HIP(hipMemcpyAsync(rf.X(), X, bytes,
                        ptr_kind == NH_PTR_HOST ? hipMemcpyHostToDevice
                                                : hipMemcpyDeviceToDevice,
                        ctx->stream));
  std::vector<double> target(eigenvalues, eigenvalues + k), resid;
  int mv = 0, r = 0;
  bool done = false;
  while (!done && r < o->max_rounds) {
    NH_TRY(rf.round(target, resid, mv));
    ++r;
    done = true;
    for (int j = 0; j < k; ++j)
      if (!(resid[j] <= o->tolerance * o->hnorm)) done = false;
  }
  NH_HIP(hipMemcpyAsync(X, rf.X(), bytes,
                        ptr_kind == NH_PTR_HOST ? hipMemcpyDeviceToHost
                                                : hipMemcpyDeviceToDevice,
                        ctx->stream));
  NH_HIP(hipStreamSynchronize(ctx->stream));
  for (int j = 0; j < k; ++j) {
    eigenvalues[j] = target[j];
    residuals[j] = resid[j];
  }
  if (rounds) *rounds = r;
  if (matvecs) *matvecs = mv;
  return done ? NH_OK : NH_ERR_NOCONV;
}
