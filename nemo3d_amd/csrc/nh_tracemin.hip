// PQ-TRACEMIN: trace minimisation for the eigenvalues of H nearest a
// reference energy sigma, as the smallest eigenvalues of A = (H - sigma)^2
// (qtracemin_driver, NEMO_3D/nemo3d/src/base/qtracemin_driver.c:9-947;
// ztracemin / modified_conjugate_gradient,
// NEMO_3D/TRACEMIN/Src/pqtracemin_functions.f90:182-...).
//
// Outer step k (Sameh & Tong, J. Comput. Appl. Math. 123 (2000)):
//   orthonormalise Y (Cholesky QR twice); W = A Y; Ritz pairs of Y^H W
//   (ascending theta); X = Y V, AX = W V;
//   H-residuals ||H x_j - lambda_j x_j|| of the wanted columns decide
//   convergence (lambda_j = x_j^H H x_j, FP64);
//   projected CG on (P A P) d_j = P A x_j = A x_j - theta_j x_j with
//   P = I - X X^H, per column, stopping at ||r_j|| <= min(cgtol,
//   theta_j / theta_s) ||r_j0||;  Y = X - D.
// All H·V products go through the library's block operator (stored FP64
// or FP32 values); everything else is FP64.
#include <chrono>

#include "nh_blockla.h"
#include "nh_vector.h"

namespace nh {
namespace {

class Tracemin {
 public:
  explicit Tracemin(nh_context *c) : c_(c) {}
  ~Tracemin() {
    for (void *q : mem_) (void)hipFree(q);
  }
  nh_status run(const nh_tracemin_opts *o, nh_eig_result *res);

 private:
  nh_status alloc(double2 **q, int cols) {
    const size_t bytes = (size_t)cols * (size_t)n_ * sizeof(double2);
    void *p = nullptr;
    NH_TRY(device_malloc(c_, (void **)&p, bytes));
    NH_CHECK(p, NH_ERR_NOMEM, "TRACEMIN: allocation of %zu bytes failed", bytes);
    mem_.push_back(p);
    *q = static_cast<double2 *>(p);
    return NH_OK;
  }
  int grid() const {
    int64_t g = (n_ + 255) / 256;
    return (int)(g > 16384 ? 16384 : g);
  }
  // Y = (H - sigma)^2 X  (two block products)
  nh_status applyA(const double2 *X, double2 *Y) {
    NH_TRY(launch_hv(c_, sigma_, reinterpret_cast<const nh_cplx *>(X),
                     reinterpret_cast<nh_cplx *>(T1_), s_, n_));
    NH_TRY(launch_hv(c_, sigma_, reinterpret_cast<const nh_cplx *>(T1_),
                     reinterpret_cast<nh_cplx *>(Y), s_, n_));
    matvecs_ += 2 * s_;
    return NH_OK;
  }
  nh_status gram(const double2 *X, const double2 *Y, std::vector<cd> &G) {
    constexpr int P = 8;
    G.assign(s_ * s_, 0.0);
    std::vector<double> h(2 * P * P);
    for (int a0 = 0; a0 < s_; a0 += P)
      for (int b0 = 0; b0 < s_; b0 += P) {
        const int qa = std::min(P, s_ - a0), qb = std::min(P, s_ - b0);
        hipLaunchKernelGGL((k_gram<P>), dim3(kParts), dim3(256), 0, c_->stream,
                           n_, qa, qb, X + (size_t)a0 * n_, n_,
                           Y + (size_t)b0 * n_, n_, part_);
        hipLaunchKernelGGL(k_sum_cols, dim3(sum_cols_grid(2 * P * P)), dim3(256), 0,
                           c_->stream, kParts, 2 * P * P, part_, sum_);
        NH_HIP(hipGetLastError());
        NH_HIP(hipMemcpyAsync(h.data(), sum_, h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, c_->stream));
        NH_HIP(hipStreamSynchronize(c_->stream));
        for (int a = 0; a < qa; ++a)
          for (int b = 0; b < qb; ++b)
            G[(a0 + a) * s_ + b0 + b] =
                cd(h[2 * (a * P + b)], h[2 * (a * P + b) + 1]);
      }
    return NH_OK;
  }
  nh_status coldot(const double2 *X, const double2 *Y, std::vector<cd> &d) {
    hipLaunchKernelGGL((k_coldot<kMaxP>), dim3(kParts), dim3(256), 0,
                       c_->stream, n_, s_, X, n_, Y, n_, part_);
    hipLaunchKernelGGL(k_sum_cols, dim3(2 * kMaxP), dim3(64), 0, c_->stream, kParts,
                       2 * kMaxP, part_, sum_);
    NH_HIP(hipGetLastError());
    std::vector<double> h(2 * kMaxP);
    NH_HIP(hipMemcpyAsync(h.data(), sum_, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c_->stream));
    NH_HIP(hipStreamSynchronize(c_->stream));
    d.resize(s_);
    for (int j = 0; j < s_; ++j) d[j] = cd(h[2 * j], h[2 * j + 1]);
    return NH_OK;
  }
  // Y = beta Y + X M (M: s x s)
  nh_status gemm(const double2 *X, double2 *Y, const std::vector<cd> &M,
                 double beta) {
    std::vector<double> h(2 * s_ * s_);
    for (int k = 0; k < s_ * s_; ++k) {
      h[2 * k] = M[k].real();
      h[2 * k + 1] = M[k].imag();
    }
    NH_HIP(hipMemcpyAsync(dM_, h.data(), h.size() * sizeof(double),
                          hipMemcpyHostToDevice, c_->stream));
    hipLaunchKernelGGL((k_block_gemm<kMaxP>), dim3(grid()), dim3(256), 0,
                       c_->stream, n_, s_, s_, X, n_, Y, n_, dM_, beta);
    NH_HIP(hipGetLastError());
    NH_HIP(hipStreamSynchronize(c_->stream));
    return NH_OK;
  }
  // Y_j = a_j X_j + b_j Y_j
  nh_status colaxpby(const double2 *X, double2 *Y, const std::vector<double> &a,
                     const std::vector<double> &b) {
    std::vector<double> h(2 * s_);
    for (int j = 0; j < s_; ++j) {
      h[2 * j] = a[j];
      h[2 * j + 1] = b[j];
    }
    NH_HIP(hipMemcpyAsync(dAB_, h.data(), h.size() * sizeof(double),
                          hipMemcpyHostToDevice, c_->stream));
    hipLaunchKernelGGL((k_col_axpby<kMaxP>), dim3(grid()), dim3(256), 0,
                       c_->stream, n_, s_, X, n_, Y, n_, dAB_);
    NH_HIP(hipGetLastError());
    NH_HIP(hipStreamSynchronize(c_->stream));
    return NH_OK;
  }
  nh_status orthonormalise(double2 *Y, double2 *tmp) {
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<cd> G;
      NH_TRY(gram(Y, Y, G));
      double tr = 0.0;
      for (int k = 0; k < s_; ++k) tr += G[k * s_ + k].real();
      for (int k = 0; k < s_; ++k) G[k * s_ + k] += 1e-28 * tr;
      NH_CHECK(cholesky_upper(s_, G), NH_ERR_NOCONV,
               "TRACEMIN: block became rank deficient");
      NH_TRY(gemm(Y, tmp, inv_upper(s_, G), 0.0));
      NH_HIP(hipMemcpyAsync(Y, tmp, (size_t)s_ * n_ * sizeof(double2),
                            hipMemcpyDeviceToDevice, c_->stream));
    }
    return NH_OK;
  }

  static constexpr int kParts = 512;
  nh_context *c_;
  std::vector<void *> mem_;
  int64_t n_ = 0;
  int s_ = 0;
  double sigma_ = 0.0;
  int matvecs_ = 0;
  double2 *T1_ = nullptr;
  double *part_ = nullptr, *sum_ = nullptr, *dM_ = nullptr, *dAB_ = nullptr;
};

nh_status Tracemin::run(const nh_tracemin_opts *o, nh_eig_result *res) {
  nh_context *c = c_;
  const auto t0 = std::chrono::steady_clock::now();
  NH_CHECK(c->layout == NH_LAYOUT_BLOCKS || sk_like(c),
           NH_ERR_STATE, "TRACEMIN: build the Hamiltonian first");
  s_ = o->block_size > 0 ? o->block_size : std::min(kMaxP, 2 * o->num_eig);
  NH_CHECK(o->num_eig >= 1 && s_ >= o->num_eig && s_ <= kMaxP, NH_ERR_ARG,
           "TRACEMIN: need 1 <= num_eig <= block_size <= %d", kMaxP);
  n_ = c->n;
  sigma_ = o->ref_eig;
  double2 *Y, *W, *X, *AX, *D, *R, *Pd, *Qc, *tmp;
  NH_TRY(alloc(&Y, s_));
  NH_TRY(alloc(&W, s_));
  NH_TRY(alloc(&X, s_));
  NH_TRY(alloc(&AX, s_));
  NH_TRY(alloc(&D, s_));
  NH_TRY(alloc(&R, s_));
  NH_TRY(alloc(&Pd, s_));
  NH_TRY(alloc(&Qc, s_));
  NH_TRY(alloc(&tmp, s_));
  NH_TRY(alloc(&T1_, s_));
  NH_TRY(device_malloc(c_, (void **)&part_, (size_t)kParts * 2 * kMaxP * kMaxP * sizeof(double)));
  mem_.push_back(part_);
  NH_TRY(device_malloc(c_, (void **)&sum_, 2 * kMaxP * kMaxP * sizeof(double)));
  mem_.push_back(sum_);
  NH_TRY(device_malloc(c_, (void **)&dM_, 2 * kMaxP * kMaxP * sizeof(double)));
  mem_.push_back(dM_);
  NH_TRY(device_malloc(c_, (void **)&dAB_, 2 * kMaxP * sizeof(double)));
  mem_.push_back(dAB_);

  // start block: the caller's, or the reference's random block (uniform
  // (0,1) real and imaginary parts from LAPACK zlarnv)
  {
    std::vector<nh_cplx> S((size_t)s_ * n_);
    if (o->start_block) {
      std::copy(o->start_block, o->start_block + (size_t)s_ * n_, S.begin());
    } else {
      // zlarnv(1, iseed, n, Y), iseed = {103 + rank, 101 rank, 7, 57}
      // (qtracemin_driver.c:105-116) with rank = seed
      const uint64_t r = o->seed;
      int32_t iseed[4] = {(int32_t)((103 + r) % 4096), (int32_t)((101 * r) % 4096), 7, 57};
      zlarnv_uniform(iseed, (int64_t)S.size(), S.data());
    }
    NH_CHECK(Y && !S.empty(), NH_ERR_STATE,
             "TRACEMIN: start block of %d x %lld not allocated", s_, (long long)n_);
    NH_HIP(hipMemcpyAsync(Y, S.data(), (size_t)s_ * n_ * sizeof(double2),
                          hipMemcpyHostToDevice, c->stream));
    NH_HIP(hipStreamSynchronize(c->stream));
  }
  const int ne = o->num_eig;
  std::vector<double> theta, lam(s_), resid(s_);
  bool done = false;
  int outer = 0;
  for (; outer < o->max_outer && !done; ++outer) {
    NH_TRY(orthonormalise(Y, tmp));
    NH_TRY(applyA(Y, W));
    std::vector<cd> Hk;
    NH_TRY(gram(Y, W, Hk));
    for (int a = 0; a < s_; ++a)
      for (int b = a; b < s_; ++b) {
        const cd h = 0.5 * (Hk[a * s_ + b] + std::conj(Hk[b * s_ + a]));
        Hk[a * s_ + b] = h;
        Hk[b * s_ + a] = std::conj(h);
      }
    std::vector<std::vector<cd>> V;
    dense_eigh(s_, Hk, theta, V);
    std::vector<cd> M(s_ * s_);
    for (int a = 0; a < s_; ++a)
      for (int j = 0; j < s_; ++j) M[a * s_ + j] = V[j][a];
    NH_TRY(gemm(Y, X, M, 0.0));
    NH_TRY(gemm(W, AX, M, 0.0));
    // H-residuals of the Ritz vectors: T1 = (H - sigma) X
    NH_TRY(launch_hv(c, sigma_, reinterpret_cast<const nh_cplx *>(X),
                     reinterpret_cast<nh_cplx *>(T1_), s_, n_));
    matvecs_ += s_;
    std::vector<cd> mu, rr;
    NH_TRY(coldot(X, T1_, mu));
    // tmp = T1 - mu X ; norms
    NH_HIP(hipMemcpyAsync(tmp, T1_, (size_t)s_ * n_ * sizeof(double2),
                          hipMemcpyDeviceToDevice, c->stream));
    {
      std::vector<double> a(s_), b(s_, 1.0);
      for (int j = 0; j < s_; ++j) a[j] = -mu[j].real();
      NH_TRY(colaxpby(X, tmp, a, b));
    }
    NH_TRY(coldot(tmp, tmp, rr));
    done = true;
    for (int j = 0; j < s_; ++j) {
      lam[j] = sigma_ + mu[j].real();
      resid[j] = std::sqrt(std::max(0.0, rr[j].real()));
      if (j < ne && resid[j] > o->tolerance * o->hnorm) done = false;
    }
    if (done) break;
    // projected CG: R = AX - X Theta (= P A X), D = 0, Pd = R
    NH_HIP(hipMemcpyAsync(R, AX, (size_t)s_ * n_ * sizeof(double2),
                          hipMemcpyDeviceToDevice, c->stream));
    {
      std::vector<double> a(s_), b(s_, 1.0);
      for (int j = 0; j < s_; ++j) a[j] = -theta[j];
      NH_TRY(colaxpby(X, R, a, b));
    }
    NH_HIP(hipMemsetAsync(D, 0, (size_t)s_ * n_ * sizeof(double2), c->stream));
    NH_HIP(hipMemcpyAsync(Pd, R, (size_t)s_ * n_ * sizeof(double2),
                          hipMemcpyDeviceToDevice, c->stream));
    std::vector<cd> r0;
    NH_TRY(coldot(R, R, r0));
    std::vector<double> rho(s_), stop(s_);
    std::vector<char> active(s_, 1);
    for (int j = 0; j < s_; ++j) {
      rho[j] = r0[j].real();
      const double ratio = theta[s_ - 1] > 0 ? theta[j] / theta[s_ - 1] : 1.0;
      stop[j] = std::min(o->cgtol, ratio) * std::sqrt(rho[j]);
    }
    for (int it = 0; it < o->cg_max; ++it) {
      bool any = false;
      for (int j = 0; j < s_; ++j) any |= active[j] != 0;
      if (!any) break;
      // Qc = P A Pd
      NH_TRY(applyA(Pd, Qc));
      std::vector<cd> XtQ;
      NH_TRY(gram(X, Qc, XtQ));
      std::vector<cd> Mneg(s_ * s_);
      for (int k = 0; k < s_ * s_; ++k) Mneg[k] = -XtQ[k];
      NH_TRY(gemm(X, Qc, Mneg, 1.0));
      std::vector<cd> pq;
      NH_TRY(coldot(Pd, Qc, pq));
      std::vector<double> al(s_, 0.0), one(s_, 1.0), mal(s_, 0.0);
      for (int j = 0; j < s_; ++j)
        if (active[j] && pq[j].real() > 0) {
          al[j] = rho[j] / pq[j].real();
          mal[j] = -al[j];
        }
      NH_TRY(colaxpby(Pd, D, al, one));
      NH_TRY(colaxpby(Qc, R, mal, one));
      std::vector<cd> rn;
      NH_TRY(coldot(R, R, rn));
      std::vector<double> a1(s_, 1.0), bet(s_, 0.0);
      for (int j = 0; j < s_; ++j) {
        if (!active[j]) {
          a1[j] = 0.0;
          bet[j] = 1.0;
          continue;
        }
        const double rnew = rn[j].real();
        bet[j] = rho[j] > 0 ? rnew / rho[j] : 0.0;
        rho[j] = rnew;
        if (std::sqrt(rnew) <= stop[j] || !(pq[j].real() > 0)) active[j] = 0;
      }
      // Pd = R + beta Pd  (inactive columns untouched)
      NH_TRY(colaxpby(R, Pd, a1, bet));
    }
    // Y = X - D
    NH_HIP(hipMemcpyAsync(Y, X, (size_t)s_ * n_ * sizeof(double2),
                          hipMemcpyDeviceToDevice, c->stream));
    {
      std::vector<double> a(s_, -1.0), b(s_, 1.0);
      NH_TRY(colaxpby(D, Y, a, b));
    }
  }
  // wanted pairs: the num_eig Ritz vectors with the smallest theta
  std::vector<int> idx(ne);
  for (int j = 0; j < ne; ++j) idx[j] = j;
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return lam[a] < lam[b]; });
  const int nout = std::min(ne, res->capacity);
  res->n_found = nout;
  res->iterations = outer;
  for (int e = 0; e < nout; ++e) {
    const int j = idx[e];
    res->eigenvalues[e] = lam[j];
    if (res->residuals) res->residuals[e] = resid[j];
    if (res->eigenvectors)
      NH_HIP(hipMemcpy(res->eigenvectors + (size_t)e * n_, X + (size_t)j * n_,
                       n_ * sizeof(nh_cplx), hipMemcpyDefault));  // host or device
  }
  res->matvecs = matvecs_;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return done ? NH_OK : NH_ERR_NOCONV;
}

}  // namespace
}  // namespace nh

extern "C" nh_status nh_eigensolve_tracemin(nh_context *ctx,
                                            const nh_tracemin_opts *opts,
                                            nh_eig_result *res) {
  NH_CHECK(ctx && opts && res && res->eigenvalues && res->capacity > 0,
           NH_ERR_ARG, "nh_eigensolve_tracemin: bad arguments");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  nh::Tracemin t(ctx);
  return t.run(opts, res);
}
