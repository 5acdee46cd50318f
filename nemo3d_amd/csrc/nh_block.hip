// Block Lanczos (eigensolve_blk_lanczos, NEMO_3D/eigen/src/lanczos_blk.c:
// 2145-...) on the device.
//
// Recurrence (single_iteration, lanczos_blk.c:392-550), block width p:
//   W = (H Q_j)  [block H·V, FP64 or FP32 storage]
//   W -= Q_{j-1} B_{j-1}^H
//   A_j = Q_j^H W ;  W -= Q_j A_j
//   W = Q_{j+1} B_j  (Cholesky QR applied twice; the reference uses a
//                     Gram-Schmidt QR, block_Lanczos_QR, lanczos_blk.c:273-357)
// The block tridiagonal T (diagonal blocks A_j, sub-diagonal B_j) is a
// Hermitian band matrix of bandwidth p; its eigenvalues in the Erange
// windows come from Sturm (LDL^H inertia) bisection on the band and their
// residual estimates ||B_m s_m|| from band inverse iteration (the reference
// calls zhbevx on the band, solve_block_tridiagonal_eigenvalues,
// lanczos_blk.c:891-991).  Eigenvectors come from a second pass over the
// stored A_j, B_j.  All vector arithmetic is FP64.
//
// Deflation (DeflationOn / DefTolerance, deflation_test, lanczos_blk.c:
// 225-268): when a column of W (or of the start block) keeps a norm below
// the tolerance after orthogonalisation against the columns before it
// (the diagonal of its Cholesky factor), it is dropped and the block
// shrinks for the rest of the run; B_j then holds the dropped column's
// coefficients on the new block, restricted to the rows of the columns
// before it (the rest of that column is below the tolerance).  The
// reference drops at most one column per step; here every such column of
// the step goes, one at a time (a start block may hold a level twice).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>

#include "nh_blockla.h"
#include "nh_internal.h"
#include "nh_vector.h"

namespace nh {
namespace {

struct Pair {
  double value;
  int block;     // blocks used when it converged
  int window;
};

class BlockLanczos {
 public:
  explicit BlockLanczos(nh_context *c) : c_(c) {}
  ~BlockLanczos() {
    for (void *q : mem_) (void)hipFree(q);
  }
  nh_status run(const nh_blanczos_opts *o, nh_eig_result *res);

 private:
  nh_status alloc(void **q, size_t bytes) {
    NH_TRY(device_malloc(c_, (void **)q, bytes));
    mem_.push_back(*q);
    return NH_OK;
  }
  nh_status gram(const double2 *X, int qa, const double2 *Y, int qb,
                 std::vector<cd> &G);
  nh_status gemm(const double2 *X, int qx, double2 *Y, int qy,
                 const std::vector<cd> &M, double beta);
  struct Factors {
    std::vector<cd> R1, R2;
    std::vector<int> drop;  // deflation: positions dropped, in order
  };
  nh_status drop_column(double2 *W, int q, int i, int slot);
  nh_status orthonormalise(double2 *&W, double2 *&Qout, int q, std::vector<cd> &R,
                           Factors *f, int &qn);
  nh_status replay_qr(double2 *&W, double2 *&Qout, int q, const Factors &f);
  nh_status step(int j, bool first_pass);
  void analyse(const nh_blanczos_opts *o, double hnorm);
  bool window_done(int w, int req, int rule) const;

  nh_context *c_;
  std::vector<void *> mem_;
  int p_ = 0;                // block width of the start block
  std::vector<int> sz_;      // size of block j (shrinks by deflation)
  bool deflate_ = false;
  double dtol_ = 0.0;
  double2 *Wd_ = nullptr;    // the columns a deflation drops
  int64_t n_ = 0;
  double2 *Qprev_ = nullptr, *Q_ = nullptr, *W_ = nullptr;
  double *dM_ = nullptr, *part_ = nullptr, *sum_ = nullptr;
  int nparts_ = 0;
  BandT T_;
  std::vector<cd> Bnext_;
  std::vector<Factors> qr_;  // Cholesky QR factors of every step
  KernelTimer timer_;        // sampled block H·V launches
  std::vector<Pair> conv_;
  std::vector<double> lastRitz_[2];
  std::vector<char> lastConv_[2];
};

nh_status BlockLanczos::gram(const double2 *X, int qa_all, const double2 *Y,
                             int qb_all, std::vector<cd> &G) {
  // 8x8 tiles of G = X^H Y (qa_all x qb_all; a larger register tile would
  // spill)
  constexpr int P = 8;
  G.assign(qa_all * qb_all, 0.0);
  std::vector<double> h(2 * P * P);
  for (int a0 = 0; a0 < qa_all; a0 += P)
    for (int b0 = 0; b0 < qb_all; b0 += P) {
      const int qa = std::min(P, qa_all - a0), qb = std::min(P, qb_all - b0);
      hipLaunchKernelGGL((k_gram<P>), dim3(nparts_), dim3(256), 0, c_->stream,
                         n_, qa, qb, X + (size_t)a0 * n_, n_,
                         Y + (size_t)b0 * n_, n_, part_);
      hipLaunchKernelGGL(k_sum_cols, dim3(sum_cols_grid(2 * P * P)), dim3(256), 0,
                         c_->stream, nparts_, 2 * P * P, part_, sum_);
      NH_HIP(hipGetLastError());
      NH_HIP(hipMemcpyAsync(h.data(), sum_, h.size() * sizeof(double),
                            hipMemcpyDeviceToHost, c_->stream));
      NH_HIP(hipStreamSynchronize(c_->stream));
      for (int a = 0; a < qa; ++a)
        for (int b = 0; b < qb; ++b)
          G[(a0 + a) * qb_all + b0 + b] =
              cd(h[2 * (a * P + b)], h[2 * (a * P + b) + 1]);
    }
  return NH_OK;
}

nh_status BlockLanczos::gemm(const double2 *X, int qx, double2 *Y, int qy,
                             const std::vector<cd> &M, double beta) {
  std::vector<double> h(2 * qx * qy);
  for (int k = 0; k < qx * qy; ++k) {
    h[2 * k] = M[k].real();
    h[2 * k + 1] = M[k].imag();
  }
  NH_HIP(hipMemcpyAsync(dM_, h.data(), h.size() * sizeof(double),
                        hipMemcpyHostToDevice, c_->stream));
  int64_t g = (n_ + 255) / 256;
  if (g > 16384) g = 16384;
  const int P = std::max(qx, qy) <= 4 ? 4 : std::max(qx, qy) <= 8 ? 8 : 16;
#define GM(PP)                                                                  \
  hipLaunchKernelGGL((k_block_gemm<PP>), dim3((int)g), dim3(256), 0, c_->stream, \
                     n_, qx, qy, X, n_, Y, n_, dM_, beta)
  if (P == 4) GM(4);
  else if (P == 8) GM(8);
  else GM(16);
#undef GM
  NH_HIP(hipGetLastError());
  // the host array h is released on return: wait for the copy
  NH_HIP(hipStreamSynchronize(c_->stream));
  return NH_OK;
}

// Removes column i of the q-column block W (columns after it move down by
// one); the removed column goes to slot `slot` of Wd_.
nh_status BlockLanczos::drop_column(double2 *W, int q, int i, int slot) {
  NH_HIP(hipMemcpyAsync(Wd_ + (size_t)slot * n_, W + (size_t)i * n_,
                        n_ * sizeof(double2), hipMemcpyDeviceToDevice, c_->stream));
  for (int k = i + 1; k < q; ++k)
    NH_HIP(hipMemcpyAsync(W + (size_t)(k - 1) * n_, W + (size_t)k * n_,
                          n_ * sizeof(double2), hipMemcpyDeviceToDevice,
                          c_->stream));
  return NH_OK;
}

// Qout R = W by Cholesky QR twice: Q1 = W R1^-1, Q = Q1 R2^-1, R = R2 R1
// (upper triangular).  The factors are kept so that the eigenvector pass
// repeats exactly these two products (a single product with R^-1 differs in
// rounding, and the difference grows along the recurrence).  The result ends
// in Qout, W holds Q1 afterwards (the pointers are swapped, no product is
// done in place).  With deflation on, columns whose Cholesky diagonal (the
// norm after orthogonalisation against the columns before them) is below
// the tolerance are dropped first, one at a time: qn < q columns come out,
// and R (qn x q) carries each dropped column's coefficients on the kept
// columns before it.
nh_status BlockLanczos::orthonormalise(double2 *&W, double2 *&Qout, int q,
                                       std::vector<cd> &R, Factors *f, int &qn) {
  std::vector<int> drop, dropped_orig, kept(q);
  for (int k = 0; k < q; ++k) kept[k] = k;
  if (deflate_) {
    while ((int)kept.size() > 1) {
      const int qc = (int)kept.size();
      std::vector<cd> G;
      NH_TRY(gram(W, qc, W, qc, G));
      const int d = cholesky_first_small(qc, G, dtol_);
      if (d < 0) break;
      NH_TRY(drop_column(W, qc, d, (int)drop.size()));
      drop.push_back(d);
      dropped_orig.push_back(kept[d]);
      kept.erase(kept.begin() + d);
    }
  }
  qn = (int)kept.size();
  std::vector<cd> G[2];
  for (int pass = 0; pass < 2; ++pass) {
    double2 *src = pass == 0 ? W : Qout;
    double2 *dst = pass == 0 ? Qout : W;
    NH_TRY(gram(src, qn, src, qn, G[pass]));
    // tiny diagonal lift keeps Cholesky alive for a nearly dependent block
    double tr = 0.0;
    for (int k = 0; k < qn; ++k) tr += G[pass][k * qn + k].real();
    for (int k = 0; k < qn; ++k) G[pass][k * qn + k] += 1e-28 * tr;
    NH_CHECK(cholesky_upper(qn, G[pass]), NH_ERR_NOCONV,
             "block Lanczos: block became rank deficient (deflation %s)",
             deflate_ ? "on" : "off");
    NH_TRY(gemm(src, qn, dst, qn, inv_upper(qn, G[pass]), 0.0));
  }
  std::swap(W, Qout);  // result (written to W by the second pass) -> Qout
  const std::vector<cd> Rk = matmul(qn, G[1], G[0]);
  if (drop.empty()) {
    R = Rk;
  } else {
    // coefficients of the dropped columns on the new block, kept on the
    // rows of the kept columns before each
    const int nd = (int)drop.size();
    std::vector<cd> cf;
    NH_TRY(gram(Qout, qn, Wd_, nd, cf));
    R.assign((size_t)qn * q, 0.0);
    for (int k = 0; k < qn; ++k)
      for (int r = 0; r < qn; ++r) R[r * q + kept[k]] = Rk[r * qn + k];
    for (int t = 0; t < nd; ++t) {
      const int c = dropped_orig[t];
      for (int r = 0; r < qn && kept[r] < c; ++r) R[r * q + c] = cf[r * nd + t];
    }
  }
  if (f) {
    f->R1 = G[0];
    f->R2 = G[1];
    f->drop = drop;
  }
  return NH_OK;
}

// Replays orthonormalise with stored factors: the same columns dropped,
// then Qout = (W R1^-1) R2^-1.
nh_status BlockLanczos::replay_qr(double2 *&W, double2 *&Qout, int q,
                                  const Factors &f) {
  int qc = q;
  for (size_t t = 0; t < f.drop.size(); ++t)
    NH_TRY(drop_column(W, qc--, f.drop[t], (int)t));
  NH_TRY(gemm(W, qc, Qout, qc, inv_upper(qc, f.R1), 0.0));
  NH_TRY(gemm(Qout, qc, W, qc, inv_upper(qc, f.R2), 0.0));
  std::swap(W, Qout);
  return NH_OK;
}

nh_status BlockLanczos::step(int j, bool first_pass) {
  const int p = sz_[j];
  // W = H Q_j (a sample of the first-pass products is timed)
  const bool timed = first_pass && j % 4 == 1;
  if (timed) NH_TRY(timer_.begin(c_->stream));
  NH_TRY(launch_hv(c_, 0.0, reinterpret_cast<const nh_cplx *>(Q_),
                   reinterpret_cast<nh_cplx *>(W_), p, n_));
  if (timed) NH_TRY(timer_.end(c_->stream));
  if (j > 0) {  // W -= Q_{j-1} B_{j-1}^H  (B_{j-1}: p x pm)
    const int pm = sz_[j - 1];
    const auto &B = T_.B[j - 1];
    std::vector<cd> M(pm * p);
    for (int k = 0; k < pm; ++k)
      for (int c = 0; c < p; ++c) M[k * p + c] = -std::conj(B[c * pm + k]);
    NH_TRY(gemm(Qprev_, pm, W_, p, M, 1.0));
  }
  std::vector<cd> A;
  if (first_pass) {
    NH_TRY(gram(Q_, p, W_, p, A));
    for (int a = 0; a < p; ++a)  // Hermitian part
      for (int b = a; b < p; ++b) {
        const cd h = 0.5 * (A[a * p + b] + std::conj(A[b * p + a]));
        A[a * p + b] = h;
        A[b * p + a] = std::conj(h);
      }
    T_.A.push_back(A);
  } else {
    A = T_.A[j];
  }
  std::vector<cd> M(p * p);
  for (int k = 0; k < p * p; ++k) M[k] = -A[k];
  NH_TRY(gemm(Q_, p, W_, p, M, 1.0));
  // Qprev <- new block (the old Q_{j-1} is no longer needed)
  if (first_pass) {
    Factors f;
    std::vector<cd> R;
    int pn = p;
    NH_TRY(orthonormalise(W_, Qprev_, p, R, &f, pn));
    T_.B.push_back(R);
    qr_.push_back(f);
    sz_.push_back(pn);
  } else {
    NH_TRY(replay_qr(W_, Qprev_, p, qr_[j]));
  }
  std::swap(Qprev_, Q_);  // Q_ = Q_{j+1}, Qprev_ = Q_j
  return NH_OK;
}

void BlockLanczos::analyse(const nh_blanczos_opts *o, double hnorm) {
  T_.m = (int)T_.A.size();
  T_.p = p_;
  T_.sz.assign(sz_.begin(), sz_.begin() + T_.m);
  T_.layout();
  const double tol = o->conv_tolerance * hnorm;
  // copies of one level closer than this count once: FP64 products keep
  // Kramers partners degenerate to ~1e-12, FP32 products split them by
  // ~1e-9 relative (a loose tolerance signals the FP32 stage)
  const double resolution = std::max(1e-9, 0.1 * o->conv_tolerance) * hnorm;
  std::vector<cd> work, x;
  const int req[2] = {o->num_eig_vb, o->num_eig_cb};
  const double *win[2] = {o->erange_vb, o->erange_cb};
  const auto &Bm = T_.B.back();  // sz_[m] x sz_[m-1]
  const int pl = sz_[T_.m - 1], pn = sz_[T_.m], n = T_.n();
  for (int w = 0; w < 2; ++w) {
    lastRitz_[w].clear();
    lastConv_[w].clear();
    if (req[w] <= 0) continue;
    const double lo = win[w][0], hi = win[w][1];
    const int klo = band_count(T_, lo, work), khi = band_count(T_, hi, work);
    // candidates nearest the gap; every level may come as a degenerate pair
    const int want = 2 * req[w] + 8;
    int k0 = klo, k1 = khi;
    if (w == 0) k0 = std::max(klo, khi - want);
    else k1 = std::min(khi, klo + want);
    double prev = 0.0;
    bool have_prev = false;
    for (int k = k0; k < k1; ++k) {
      double a = lo, b = hi;
      for (int it = 0; it < 200 && b - a > 1e-13 * (1.0 + std::fabs(a)); ++it) {
        const double mid = 0.5 * (a + b);
        if (band_count(T_, mid, work) > k) b = mid;
        else a = mid;
      }
      const double ev = 0.5 * (a + b);
      // copies of one level (Kramers partners, converged duplicates) count once
      if (have_prev && std::fabs(ev - prev) < resolution) continue;
      prev = ev;
      have_prev = true;
      band_eigvec(T_, ev, x);
      // residual estimate ||B_m s_m|| (s_m: last block of the Ritz vector)
      double r2 = 0.0;
      for (int rr = 0; rr < pn; ++rr) {
        cd sm = 0.0;
        for (int cc = 0; cc < pl; ++cc) sm += Bm[rr * pl + cc] * x[n - pl + cc];
        r2 += std::norm(sm);
      }
      const bool ok = std::sqrt(r2) <= tol;
      lastRitz_[w].push_back(ev);
      lastConv_[w].push_back(ok);
      if (ok) {
        bool known = false;
        for (auto &cv : conv_)
          if (std::fabs(cv.value - ev) < resolution && cv.window == w) {
            known = true;
            cv.value = ev;  // keep the latest (most accurate) Ritz value
          }
        if (!known) conv_.push_back({ev, T_.m, w});
      }
    }
  }
}

bool BlockLanczos::window_done(int w, int req, int rule) const {
  if (req <= 0) return true;
  if (rule == 0) {
    int cnt = 0;
    for (const auto &cv : conv_) cnt += cv.window == w;
    return cnt >= req;
  }
  const auto &r = lastRitz_[w];
  if ((int)r.size() < req) return false;
  for (int t = 0; t < req; ++t)
    if (!lastConv_[w][w == 0 ? r.size() - 1 - t : t]) return false;
  return true;
}

nh_status BlockLanczos::run(const nh_blanczos_opts *o, nh_eig_result *res) {
  nh_context *c = c_;
  const auto t0 = std::chrono::steady_clock::now();
  p_ = o->block_size;
  NH_CHECK(p_ >= 1 && p_ <= kMaxP, NH_ERR_ARG, "block size must be 1..%d", kMaxP);
  NH_CHECK(c->layout == NH_LAYOUT_BLOCKS || sk_like(c),
           NH_ERR_STATE, "block Lanczos: build the Hamiltonian first");
  n_ = c->n;
  const size_t blk = (size_t)p_ * n_ * sizeof(double2);
  NH_TRY(alloc((void **)&Qprev_, blk));
  NH_TRY(alloc((void **)&Q_, blk));
  NH_TRY(alloc((void **)&W_, blk));
  NH_TRY(alloc((void **)&dM_, 2 * kMaxP * kMaxP * sizeof(double)));
  nparts_ = 1024;
  NH_TRY(alloc((void **)&part_, (size_t)nparts_ * 2 * kMaxP * kMaxP * sizeof(double)));
  NH_TRY(alloc((void **)&sum_, 2 * kMaxP * kMaxP * sizeof(double)));

  // start block: the caller's block, or the reference's random block: p
  // vectors (rand() / RAND_MAX, 0) in turn after srand(seed)
  // (lanczos_blk.c:2214-2229)
  std::vector<nh_cplx> S((size_t)p_ * n_);
  if (o->start_block)
    std::copy(o->start_block, o->start_block + (size_t)p_ * n_, S.begin());
  else
    rand_start((uint32_t)o->seed, (int64_t)p_ * n_, S.data());
  NH_HIP(hipMemcpyAsync(W_, S.data(), blk, hipMemcpyHostToDevice, c->stream));
  deflate_ = o->deflation_on != 0;
  dtol_ = o->deflation_tol;
  NH_CHECK(!deflate_ || dtol_ >= 0.0, NH_ERR_ARG, "deflation_tol must be >= 0");
  if (deflate_) NH_TRY(alloc((void **)&Wd_, (size_t)p_ * n_ * sizeof(double2)));
  // the start block (with deflation on, columns that depend on the ones
  // before them are dropped: a learned block may hold a level twice)
  std::vector<cd> R0;
  int qs = p_;
  NH_TRY(orthonormalise(W_, Q_, p_, R0, nullptr, qs));
  sz_.assign(1, qs);
  // keep Q_1 for the eigenvector pass
  double2 *Q1 = nullptr;
  NH_TRY(alloc((void **)&Q1, blk));
  NH_HIP(hipMemcpyAsync(Q1, Q_, blk, hipMemcpyDeviceToDevice, c->stream));

  int j = 0, skip = 0, matvecs = 0;
  bool done = false;
  double hnorm = 1.0;
  Progress prog;
  while (!done && j < o->max_iter) {
    NH_TRY(step(j, true));
    matvecs += sz_[j];
    ++j;
    prog.tick("block lanczos", j, (int)conv_.size());
    ++skip;
    if (j > o->conv_check_start && skip >= o->conv_check_skip) {
      skip = 0;
      // ||H|| estimate from the extreme Ritz values (Gershgorin on T)
      T_.m = j;
      T_.p = p_;
      T_.sz.assign(sz_.begin(), sz_.begin() + j);
      T_.layout();
      double g = 0.0;
      for (int r = 0; r < T_.n(); ++r) {
        double s = 0.0;
        for (int cc = std::max(0, r - p_); cc <= std::min(T_.n() - 1, r + p_); ++cc)
          s += std::abs(cc >= r ? T_.at(r, cc) : std::conj(T_.at(cc, r)));
        g = std::max(g, s);
      }
      hnorm = g;
      analyse(o, hnorm);
      done = window_done(0, o->num_eig_vb, o->stop_rule) &&
             window_done(1, o->num_eig_cb, o->stop_rule);
    }
  }
  if (!done) {
    analyse(o, hnorm);
    done = window_done(0, o->num_eig_vb, o->stop_rule) &&
           window_done(1, o->num_eig_cb, o->stop_rule);
  }

  std::vector<double> pick;
  const int req[2] = {o->num_eig_vb, o->num_eig_cb};
  for (int w = 0; w < 2; ++w) {
    if (o->stop_rule == 0) {
      // the converged levels of the window nearest the gap
      std::vector<double> in;
      for (const auto &cv : conv_)
        if (cv.window == w) in.push_back(cv.value);
      std::sort(in.begin(), in.end());
      if (w == 0) std::reverse(in.begin(), in.end());
      for (int t = 0; t < (int)in.size() && t < req[w]; ++t) pick.push_back(in[t]);
      continue;
    }
    const auto &r = lastRitz_[w];
    for (int t = 0; t < (int)r.size() && t < req[w]; ++t) {
      const int idx = w == 0 ? (int)r.size() - 1 - t : t;
      if (!lastConv_[w][idx]) break;
      pick.push_back(r[idx]);
    }
  }
  std::sort(pick.begin(), pick.end());
  const int ne = std::min((int)pick.size(), res->capacity);
  res->n_found = ne;
  res->iterations = j;
  for (int e = 0; e < ne; ++e) {
    res->eigenvalues[e] = pick[e];
    if (res->residuals) res->residuals[e] = -1.0;
  }
  if (o->want_vectors && ne > 0) {
    // eigenvectors of T (all j blocks), then a second pass accumulating
    // V_e += Q_k s_e[k]
    T_.m = j;
    T_.sz.assign(sz_.begin(), sz_.begin() + j);
    T_.layout();
    std::vector<std::vector<cd>> s(ne);
    for (int e = 0; e < ne; ++e) {
      std::vector<const std::vector<cd> *> cluster;
      for (int f = 0; f < e; ++f)
        if (std::fabs(pick[f] - pick[e]) <= 1e-9 * (1.0 + std::fabs(pick[e])))
          cluster.push_back(&s[f]);
      band_eigvec(T_, pick[e], s[e], (uint64_t)e, cluster);
    }
    double2 *V = nullptr;
    NH_TRY(alloc((void **)&V, (size_t)ne * n_ * sizeof(double2)));
    NH_HIP(hipMemsetAsync(V, 0, (size_t)ne * n_ * sizeof(double2), c->stream));
    NH_HIP(hipMemcpyAsync(Q_, Q1, blk, hipMemcpyDeviceToDevice, c->stream));
    const int steps = j;
    for (int k = 0; k < steps; ++k) {
      // V[:, e] += Q_k s_e[off_k .. off_k + sz_k)
      const int pk = sz_[k], ok = T_.offset(k);
      for (int e0 = 0; e0 < ne; e0 += kMaxP) {
        const int qe = std::min(kMaxP, ne - e0);
        std::vector<cd> M(pk * qe);
        for (int a = 0; a < pk; ++a)
          for (int e = 0; e < qe; ++e) M[a * qe + e] = s[e0 + e][ok + a];
        // gemm writes qe columns with leading dimension n_
        std::vector<double> h(2 * pk * qe);
        for (int t = 0; t < pk * qe; ++t) {
          h[2 * t] = M[t].real();
          h[2 * t + 1] = M[t].imag();
        }
        NH_HIP(hipMemcpyAsync(dM_, h.data(), h.size() * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
        int64_t g = (n_ + 255) / 256;
        if (g > 16384) g = 16384;
        hipLaunchKernelGGL((k_block_gemm<kMaxP>), dim3((int)g), dim3(256), 0,
                           c->stream, n_, pk, qe, Q_, n_, V + (size_t)e0 * n_,
                           n_, dM_, 1.0);
        NH_HIP(hipGetLastError());
        NH_HIP(hipStreamSynchronize(c->stream));
      }
      if (k + 1 < steps) NH_TRY(step(k, false));
      matvecs += pk;
    }
    // normalise, Rayleigh quotient, FP64 residual
    double2 *Y = W_;
    for (int e = 0; e < ne; ++e) {
      const nh_cplx *v = reinterpret_cast<const nh_cplx *>(V + (size_t)e * n_);
      double nrm2;
      NH_TRY(reduce_nrm2sq(c, n_, v, &nrm2));
      const nh_cplx sc = {1.0 / std::sqrt(nrm2), 0.0}, z = {0.0, 0.0};
      NH_TRY(launch_axpby(c, n_, sc, v, z, reinterpret_cast<nh_cplx *>(V + (size_t)e * n_)));
      NH_TRY(launch_hv(c, 0.0, v, reinterpret_cast<nh_cplx *>(Y), 1, n_));
      ++matvecs;
      double rq[2];
      NH_TRY(reduce_dot(c, n_, v, reinterpret_cast<nh_cplx *>(Y), rq));
      const nh_cplx a = {-rq[0], 0.0}, b = {1.0, 0.0};
      NH_TRY(launch_axpby(c, n_, a, v, b, reinterpret_cast<nh_cplx *>(Y)));
      double r2;
      NH_TRY(reduce_nrm2sq(c, n_, reinterpret_cast<nh_cplx *>(Y), &r2));
      res->eigenvalues[e] = rq[0];
      if (res->residuals) res->residuals[e] = std::sqrt(r2);
      if (res->eigenvectors)
        NH_HIP(hipMemcpy(res->eigenvectors + (size_t)e * n_, v,
                         n_ * sizeof(nh_cplx), hipMemcpyDefault));  // host or device
    }
  }
  NH_HIP(hipStreamSynchronize(c->stream));
  res->matvecs = matvecs;
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  NH_TRY(timer_.collect(c->stream));
  res->hv_ms = timer_.mean_ms();
  res->hv_timed = timer_.count;
  return done ? NH_OK : NH_ERR_NOCONV;
}

}  // namespace
}  // namespace nh

extern "C" nh_status nh_eigensolve_block_lanczos(nh_context *ctx,
                                                 const nh_blanczos_opts *opts,
                                                 nh_eig_result *res) {
  NH_CHECK(ctx && opts && res && res->eigenvalues && res->capacity > 0,
           NH_ERR_ARG, "nh_eigensolve_block_lanczos: bad arguments");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  nh::BlockLanczos b(ctx);
  return b.run(opts, res);
}
