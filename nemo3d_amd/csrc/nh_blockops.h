// Host-driven block operations on N-length device vectors stored column
// after column ([col][N]): Gram X^H Y, update Y = beta Y + X M and Cholesky QR
// applied twice.  Small matrices live on the host (row-major, complex).  A
// Gram is one tiled launch plus one reduction; an update streams all the
// columns of X in one pass, so a Gram or update against a wide basis reads
// the basis once.
#pragma once

#include <algorithm>
#include <vector>

#include "nh_blockla.h"
#include "nh_internal.h"
#include "nh_vector.h"

namespace nh {

class BlockOps {
 public:
  static constexpr int kTile = 8;     // Gram tile (k_gram_tiles<8>)
  static constexpr int kChunks = 512;  // row chunks of a Gram

  // max_cols: widest block passed to cholqr and widest Y of a Gram or an
  // update (<= 16); max_x: most columns of X in a Gram or an update
  // (default max_cols)
  BlockOps(nh_context *c, int max_cols, int max_x = 0)
      : c_(c), n_(c->n), cap_(max_cols), maxx_(std::max(max_x, max_cols)) {}
  ~BlockOps() {
    for (void *q : mem_) pool_release(c_, q);
  }
  nh_status setup() {
    NH_CHECK(cap_ >= 1 && cap_ <= 16, NH_ERR_ARG, "block width %d", cap_);
    chunk_ = std::max<int64_t>(256, ((n_ + kChunks - 1) / kChunks + 255) / 256 * 256);
    nch_ = (int)((n_ + chunk_ - 1) / chunk_);
    maxt_ = ((maxx_ + kTile - 1) / kTile) * ((cap_ + kTile - 1) / kTile);
    const int len = maxt_ * 2 * kTile * kTile;
    NH_TRY(alloc((void **)&part_, (size_t)nch_ * len * sizeof(double)));
    NH_TRY(alloc((void **)&sum_, (size_t)len * sizeof(double)));
    NH_TRY(alloc((void **)&dM_, (size_t)maxx_ * 16 * sizeof(double2)));
    const int capp = (cap_ + kTile - 1) / kTile * kTile;
    NH_TRY(alloc((void **)&S_, (size_t)capp * n_ * sizeof(double2)));
    return NH_OK;
  }
  // device memory owned by this object (from the context's workspace
  // pool), zero-filled
  nh_status alloc(void **q, size_t bytes) {
    NH_TRY(pool_alloc(c_, q, bytes));
    mem_.push_back(*q);
    NH_HIP(hipMemsetAsync(*q, 0, bytes, c_->stream));
    return NH_OK;
  }

  // G (qx x qy, row-major) = X^H Y; qx <= max_x, qy <= max_cols.  One
  // launch covers every 8 x 8 tile; one host synchronisation.
  nh_status gram(const double2 *X, int qx, const double2 *Y, int qy,
                 std::vector<cd> &G) {
    NH_CHECK(qx >= 1 && qx <= maxx_ && qy >= 1 && qy <= cap_, NH_ERR_ARG,
             "gram shape %d x %d", qx, qy);
    const int ntx = (qx + kTile - 1) / kTile, nty = (qy + kTile - 1) / kTile;
    const int nt = ntx * nty, len = nt * 2 * kTile * kTile;
    hipLaunchKernelGGL((k_gram_tiles<kTile>), dim3(nch_ * nt), dim3(256), 0,
                       c_->stream, n_, chunk_, qx, qy, X, n_, Y, n_, part_);
    hipLaunchKernelGGL(k_sum_cols, dim3(sum_cols_grid(len)), dim3(256), 0,
                       c_->stream, nch_, len, part_, sum_);
    NH_HIP(hipGetLastError());
    std::vector<double> h(len);
    NH_HIP(hipMemcpyAsync(h.data(), sum_, h.size() * sizeof(double),
                          hipMemcpyDeviceToHost, c_->stream));
    NH_HIP(hipStreamSynchronize(c_->stream));
    G.assign((size_t)qx * qy, 0.0);
    for (int t = 0; t < nt; ++t) {
      const int tx = t / nty, ty = t % nty;
      const double *e = h.data() + (size_t)t * 2 * kTile * kTile;
      for (int a = 0; a < kTile && tx * kTile + a < qx; ++a)
        for (int b = 0; b < kTile && ty * kTile + b < qy; ++b)
          G[(size_t)(tx * kTile + a) * qy + ty * kTile + b] =
              cd(e[2 * (a * kTile + b)], e[2 * (a * kTile + b) + 1]);
    }
    return NH_OK;
  }

  // Y[:, 0..qy) = beta Y + X M with M (qx x qy, row-major), qx <= max_x,
  // qy <= 16; Y must not overlap X.  One pass over the rows per
  // kGemmMaxX columns of X.
  nh_status gemm(const double2 *X, int qx, double2 *Y, int qy,
                 const std::vector<cd> &M, double beta) {
    NH_CHECK(qy >= 1 && qy <= 16 && qx >= 1 && qx <= maxx_, NH_ERR_ARG,
             "block update shape %d x %d", qx, qy);
    static_assert(sizeof(cd) == sizeof(double2), "complex layout");
    NH_HIP(hipMemcpyAsync(dM_, M.data(), (size_t)qx * qy * sizeof(double2),
                          hipMemcpyHostToDevice, c_->stream));
    int64_t g = (n_ + 255) / 256;
    if (g > 16384) g = 16384;
    for (int x0 = 0; x0 < qx; x0 += kGemmMaxX) {
      const int q = std::min(kGemmMaxX, qx - x0);
      const double b = x0 == 0 ? beta : 1.0;
      if (qy <= 8)
        hipLaunchKernelGGL((k_gemm_wide<8>), dim3((int)g), dim3(256), 0,
                           c_->stream, n_, q, qy, X + (size_t)x0 * n_, n_, Y, n_,
                           dM_ + (size_t)x0 * qy, b);
      else
        hipLaunchKernelGGL((k_gemm_wide<16>), dim3((int)g), dim3(256), 0,
                           c_->stream, n_, q, qy, X + (size_t)x0 * n_, n_, Y, n_,
                           dM_ + (size_t)x0 * qy, b);
      NH_HIP(hipGetLastError());
    }
    // M is the caller's: keep it until the copy has completed
    NH_HIP(hipStreamSynchronize(c_->stream));
    return NH_OK;
  }

  // Q R = W (k columns, k <= max_cols) by Cholesky QR applied twice; R =
  // R2 R1.  ok = false when W is numerically rank deficient (Q may then
  // hold partial results).  W is left unchanged, Q must not overlap W.
  nh_status cholqr(const double2 *W, double2 *Q, int k, std::vector<cd> &R,
                   bool &ok) {
    NH_CHECK(k >= 1 && k <= cap_, NH_ERR_ARG, "cholqr width %d", k);
    std::vector<cd> G, F[2];
    ok = true;
    for (int pass = 0; pass < 2; ++pass) {
      const double2 *src = pass == 0 ? W : Q;
      double2 *dst = pass == 0 ? Q : S_;
      NH_TRY(gram(src, k, src, k, G));
      F[pass] = G;
      double dmax = 0.0, dmin = 1e300;
      for (int a = 0; a < k; ++a) {
        dmax = std::max(dmax, F[pass][a * k + a].real());
        dmin = std::min(dmin, F[pass][a * k + a].real());
      }
      if (!(dmin > 1e-24 * std::max(dmax, 1e-300)) || !cholesky_upper(k, F[pass])) {
        ok = false;
        return NH_OK;
      }
      NH_TRY(gemm(src, k, dst, k, inv_upper(k, F[pass]), 0.0));
    }
    NH_HIP(hipMemcpyAsync(Q, S_, (size_t)k * n_ * sizeof(double2),
                          hipMemcpyDeviceToDevice, c_->stream));
    R = matmul(k, F[1], F[0]);
    return NH_OK;
  }

  int64_t n() const { return n_; }
  nh_context *ctx() const { return c_; }

 private:
  nh_context *c_;
  int64_t n_;
  int cap_, maxx_;
  int64_t chunk_ = 0;
  int nch_ = 0, maxt_ = 0;
  std::vector<void *> mem_;
  double *part_ = nullptr, *sum_ = nullptr;
  double2 *dM_ = nullptr, *S_ = nullptr;
};

}  // namespace nh
