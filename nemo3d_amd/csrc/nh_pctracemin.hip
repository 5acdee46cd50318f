// PC-TRACEMIN: trace minimisation on a Chebyshev map of H
// (ctracemin_driver, NEMO_3D/nemo3d/src/base/ctracemin_driver.c:10-...;
// map_chebyshev_shift and zgtracemin,
// NEMO_3D/TRACEMIN/Src/pctracemin_functions.f90:154-332).
//
// The map A = T_k(q(H)) + 2I with q(x) = 1 - gamma (x - a)(x - b) and
// gamma = min(2 / ((s - a)(s - b)), 2 / ((l - a)(l - b))) for the wanted
// interval [a, b] inside the spectrum bounds [s, l]: q maps the unwanted
// part of the spectrum into [-1, 1] (A there in [1, 3]) and the wanted
// interval above 1, where the Chebyshev polynomial T_k grows, so the
// wanted eigenvectors dominate A.  Each outer step applies A to the block
// (2k block H·V products through the three-term Chebyshev recurrence),
// orthonormalises it (Cholesky QR twice) and takes the Rayleigh-Ritz pairs
// of H on the block, with FP64 residuals ||H y - theta y|| as the test (the
// reference tests the residual of A; the H-residual is the quantity the
// caller needs).  Block vectors are [col][N] in HBM; small matrices on the
// host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>

#include "nh_blockops.h"

namespace nh {
namespace {

class PCTracemin {
 public:
  PCTracemin(nh_context *c, const nh_pctracemin_opts *o)
      : c_(c), o_(o), ops_(c, 16) {}
  nh_status run(nh_eig_result *res);

 private:
  double2 *col(double2 *b, int j) { return b + (size_t)j * n_; }
  // y <- (H - shift) x for the p columns
  nh_status hv(const double2 *x, double2 *y, double shift) {
    NH_TRY(launch_hv(c_, shift, reinterpret_cast<const nh_cplx *>(x),
                     reinterpret_cast<nh_cplx *>(y), p_, n_));
    matvecs_ += p_;
    return NH_OK;
  }
  // y <- a x + b y over the p columns (one long vector)
  nh_status axpby(double a, const double2 *x, double b, double2 *y) {
    return launch_axpby(c_, (int64_t)p_ * n_, nh_cplx{a, 0.0},
                        reinterpret_cast<const nh_cplx *>(x), nh_cplx{b, 0.0},
                        reinterpret_cast<nh_cplx *>(y));
  }
  // out <- 2 c q(H) v - c' prev (Chebyshev step, map_chebyshev_shift):
  // W2 = (H - a)(H - b) v, out = scale (v - gamma W2) - prev_coef prev
  nh_status qstep(const double2 *v, const double2 *prev, double scale,
                  double prev_coef, double2 *&out);
  nh_status filter();  // Vc <- T_k(q(H)) Y + 2 Y

  nh_context *c_;
  const nh_pctracemin_opts *o_;
  BlockOps ops_;
  int p_ = 0, pp_ = 0;
  int64_t n_ = 0;
  double gamma_ = 0.0;
  int matvecs_ = 0;
  double2 *Y_ = nullptr, *Vp_ = nullptr, *Vc_ = nullptr, *W1_ = nullptr,
          *W2_ = nullptr, *HY_ = nullptr;
  double *dab_ = nullptr, *part_ = nullptr, *sum_ = nullptr;
  int nparts_ = 0;
};

nh_status PCTracemin::qstep(const double2 *v, const double2 *prev,
                            double scale, double prev_coef, double2 *&out) {
  const double a = o_->want_range[0], b = o_->want_range[1];
  NH_TRY(hv(v, W1_, b));                 // (H - b) v
  NH_TRY(hv(W1_, W2_, a));               // (H - a)(H - b) v
  NH_TRY(axpby(scale, v, -scale * gamma_, W2_));  // scale (v - gamma W2)
  if (prev) NH_TRY(axpby(-prev_coef, prev, 1.0, W2_));
  out = W2_;
  return NH_OK;
}

nh_status PCTracemin::filter() {
  // V0 = Y, V1 = q(H) Y, V_{j+1} = 2 q(H) V_j - V_{j-1}; A Y = V_k + 2 Y
  double2 *out = nullptr;
  NH_TRY(qstep(Y_, nullptr, 1.0, 0.0, out));
  std::swap(W2_, Vc_);  // Vc = V1
  const double2 *prev = Y_;
  for (int j = 1; j < o_->cheb_degree; ++j) {
    NH_TRY(qstep(Vc_, prev, 2.0, 1.0, out));
    // rotate: Vp <- Vc, Vc <- new, W2 <- old Vp (free)
    double2 *old_vp = Vp_;
    Vp_ = Vc_;
    Vc_ = W2_;
    W2_ = old_vp;
    prev = Vp_;
  }
  return axpby(2.0, Y_, 1.0, Vc_);
}

nh_status PCTracemin::run(nh_eig_result *res) {
  nh_context *c = c_;
  const auto t0 = std::chrono::steady_clock::now();
  const nh_pctracemin_opts *o = o_;
  const double a = o->want_range[0], b = o->want_range[1];
  const double s = o->all_range[0], l = o->all_range[1];
  p_ = o->block_size > 0 ? o->block_size : std::min(16, o->num_eig + 4);
  NH_CHECK(o->num_eig >= 1 && p_ >= o->num_eig && p_ <= 16, NH_ERR_ARG,
           "PC-TRACEMIN: need 1 <= num_eig <= block_size <= 16");
  NH_CHECK(s < a && a < b && b < l, NH_ERR_ARG,
           "PC-TRACEMIN: need AllEigRange[0] < a < b < AllEigRange[1]");
  NH_CHECK(o->cheb_degree >= 1 && o->max_outer >= 1 && o->tolerance > 0,
           NH_ERR_ARG, "PC-TRACEMIN: bad degree / maxi / tolerance");
  // ctracemin_driver.c: gamma = min(2/((saeig-seig)(saeig-leig)),
  //                                 2/((laeig-seig)(laeig-leig)))
  gamma_ = std::min(2.0 / ((s - a) * (s - b)), 2.0 / ((l - a) * (l - b)));
  const double hnorm = o->hnorm > 0 ? o->hnorm : std::max(std::fabs(s), std::fabs(l));
  n_ = c->n;
  pp_ = (p_ + 7) / 8 * 8;
  NH_TRY(ops_.setup());
  const size_t bytes = (size_t)pp_ * n_ * sizeof(double2);
  for (double2 **q : {&Y_, &Vp_, &Vc_, &W1_, &W2_, &HY_})
    NH_TRY(ops_.alloc((void **)q, bytes));
  nparts_ = reduce_blocks(n_);
  NH_TRY(ops_.alloc((void **)&dab_, 2 * 16 * sizeof(double)));
  NH_TRY(ops_.alloc((void **)&part_, (size_t)nparts_ * 2 * 16 * sizeof(double)));
  NH_TRY(ops_.alloc((void **)&sum_, 2 * 16 * sizeof(double)));

  // start block: the caller's, or the reference's random block (uniform
  // (0,1) real and imaginary parts from LAPACK zlarnv, as for PQ-TRACEMIN)
  {
    std::vector<nh_cplx> S((size_t)p_ * n_);
    if (o->start_block) {
      std::copy(o->start_block, o->start_block + (size_t)p_ * n_, S.begin());
    } else {
      // zlarnv(1, iseed, n, Y), iseed = {103 + rank, 101 rank, 7, 57}
      // (qtracemin_driver.c:105-116) with rank = seed
      const uint64_t r = o->seed;
      int32_t iseed[4] = {(int32_t)((103 + r) % 4096), (int32_t)((101 * r) % 4096), 7, 57};
      zlarnv_uniform(iseed, (int64_t)S.size(), S.data());
    }
    NH_HIP(hipMemcpyAsync(W1_, S.data(), S.size() * sizeof(double2),
                          hipMemcpyHostToDevice, c->stream));
    bool ok = false;
    std::vector<cd> R;
    NH_TRY(ops_.cholqr(W1_, Y_, p_, R, ok));
    NH_HIP(hipMemsetAsync(W1_, 0, bytes, c->stream));
    NH_CHECK(ok, NH_ERR_ARG, "PC-TRACEMIN: start block is rank deficient");
  }

  std::vector<double> theta(p_), resid(p_);
  std::vector<int> pick;
  bool done = false;
  int it = 0;
  Progress prog;
  while (!done && it < o->max_outer) {
    ++it;
    NH_TRY(filter());
    bool ok = false;
    std::vector<cd> R;
    NH_TRY(ops_.cholqr(Vc_, Y_, p_, R, ok));
    NH_CHECK(ok, NH_ERR_NOCONV,
             "PC-TRACEMIN: filtered block lost rank (degree too high for the "
             "block size?)");
    // Rayleigh-Ritz of H on the block
    NH_TRY(hv(Y_, HY_, 0.0));
    std::vector<cd> G;
    NH_TRY(ops_.gram(Y_, pp_, HY_, pp_, G));
    std::vector<cd> T((size_t)p_ * p_);
    for (int r = 0; r < p_; ++r)
      for (int q = 0; q < p_; ++q)
        T[(size_t)r * p_ + q] = 0.5 * (G[(size_t)r * pp_ + q] + std::conj(G[(size_t)q * pp_ + r]));
    std::vector<std::vector<cd>> Z;
    dense_eigh(p_, T, theta, Z);
    std::vector<cd> M((size_t)p_ * p_);
    for (int r = 0; r < p_; ++r)
      for (int q = 0; q < p_; ++q) M[(size_t)r * p_ + q] = Z[q][r];
    NH_TRY(ops_.gemm(Y_, p_, W1_, p_, M, 0.0));   // Ritz vectors
    NH_TRY(ops_.gemm(HY_, p_, W2_, p_, M, 0.0));  // H times them
    std::swap(Y_, W1_);
    std::swap(HY_, W2_);
    // residuals ||H y - theta y|| (explicit difference)
    NH_HIP(hipMemcpyAsync(W1_, HY_, (size_t)p_ * n_ * sizeof(double2),
                          hipMemcpyDeviceToDevice, c->stream));
    std::vector<double> ab(2 * p_);
    for (int j = 0; j < p_; ++j) {
      ab[2 * j] = -theta[j];
      ab[2 * j + 1] = 1.0;
    }
    NH_HIP(hipMemcpyAsync(dab_, ab.data(), ab.size() * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
    int64_t g = (n_ + 255) / 256;
    if (g > 16384) g = 16384;
    hipLaunchKernelGGL((k_col_axpby<16>), dim3((int)g), dim3(256), 0, c->stream,
                       n_, p_, Y_, n_, W1_, n_, dab_);
    hipLaunchKernelGGL((k_coldot<16>), dim3(nparts_), dim3(256), 0, c->stream,
                       n_, p_, W1_, n_, W1_, n_, part_);
    hipLaunchKernelGGL(k_sum_cols, dim3(32), dim3(256), 0, c->stream, nparts_,
                       32, part_, sum_);
    NH_HIP(hipGetLastError());
    std::vector<double> h(32);
    NH_HIP(hipMemcpyAsync(h.data(), sum_, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
    NH_HIP(hipStreamSynchronize(c->stream));
    for (int j = 0; j < p_; ++j) resid[j] = std::sqrt(std::max(h[2 * j], 0.0));
    // wanted pairs: Ritz values in [a, b], nearest the interval centre first
    pick.clear();
    for (int j = 0; j < p_; ++j)
      if (theta[j] >= a && theta[j] <= b) pick.push_back(j);
    const double mid = 0.5 * (a + b);
    std::sort(pick.begin(), pick.end(), [&](int x, int y) {
      return std::fabs(theta[x] - mid) < std::fabs(theta[y] - mid);
    });
    if ((int)pick.size() > o->num_eig) pick.resize(o->num_eig);
    done = (int)pick.size() == o->num_eig;
    for (int j : pick)
      if (!(resid[j] <= o->tolerance * hnorm)) done = false;
    prog.tick("pctracemin", it, (int)pick.size());
  }
  std::sort(pick.begin(), pick.end(),
            [&](int x, int y) { return theta[x] < theta[y]; });
  const int nout = std::min((int)pick.size(), res->capacity);
  for (int e = 0; e < nout; ++e) {
    const int j = pick[e];
    res->eigenvalues[e] = theta[j];
    if (res->residuals) res->residuals[e] = resid[j];
    if (res->eigenvectors)
      NH_HIP(hipMemcpy(res->eigenvectors + (size_t)e * n_, col(Y_, j),
                       n_ * sizeof(nh_cplx), hipMemcpyDefault));  // host or device
  }
  res->n_found = nout;
  res->iterations = it;
  res->matvecs = matvecs_;
  res->hv_ms = 0.0;
  res->hv_timed = 0;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return done ? NH_OK : NH_ERR_NOCONV;
}

}  // namespace
}  // namespace nh

using namespace nh;

extern "C" nh_status nh_eigensolve_pctracemin(nh_context *ctx,
                                              const nh_pctracemin_opts *opts,
                                              nh_eig_result *res) {
  NH_CHECK(ctx && opts && res && res->eigenvalues && res->capacity >= 1,
           NH_ERR_ARG, "PC-TRACEMIN: null argument");
  NH_CHECK(ctx->layout == NH_LAYOUT_BLOCKS || sk_like(ctx),
           NH_ERR_STATE, "PC-TRACEMIN: build the Hamiltonian first");
  NH_HIP(hipSetDevice(ctx->device));
  res->n_found = 0;
  PCTracemin t(ctx, opts);
  return t.run(res);
}
