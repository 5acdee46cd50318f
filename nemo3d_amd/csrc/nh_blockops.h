// Host-driven block operations on N-length device vectors stored column
// after column ([col][N]): Gram X^H Y, update Y = beta Y + X M and Cholesky QR
// applied twice.  Small matrices live on the host (row-major, complex).
#pragma once

#include <algorithm>
#include <vector>

#include "nh_blockla.h"
#include "nh_internal.h"
#include "nh_vector.h"

namespace nh {

class BlockOps {
 public:
  static constexpr int kTile = 8;  // Gram tile (k_gram<8>)

  // max_cols: widest block passed to cholqr (<= 16)
  BlockOps(nh_context *c, int max_cols) : c_(c), n_(c->n), cap_(max_cols) {}
  ~BlockOps() {
    for (void *q : mem_) hipFree(q);
  }
  nh_status setup() {
    nparts_ = reduce_blocks(n_);
    NH_TRY(alloc((void **)&part_, (size_t)nparts_ * 2 * kTile * kTile * sizeof(double)));
    NH_TRY(alloc((void **)&sum_, 2 * kTile * kTile * sizeof(double)));
    NH_TRY(alloc((void **)&dM_, 2 * 16 * 16 * sizeof(double)));
    const int capp = (cap_ + kTile - 1) / kTile * kTile;
    NH_TRY(alloc((void **)&S_, (size_t)capp * n_ * sizeof(double2)));
    return NH_OK;
  }
  // device memory owned by this object, zero-filled
  nh_status alloc(void **q, size_t bytes) {
    NH_HIP(hipMalloc(q, bytes));
    mem_.push_back(*q);
    NH_HIP(hipMemsetAsync(*q, 0, bytes, c_->stream));
    return NH_OK;
  }

  // G (qx x qy) = X^H Y.  Columns are read in tiles of 8: X and Y must
  // have round_up(qx, 8) / round_up(qy, 8) columns allocated (the padding
  // columns zero where their products matter).
  nh_status gram(const double2 *X, int qx, const double2 *Y, int qy,
                 std::vector<cd> &G) {
    G.assign((size_t)qx * qy, 0.0);
    std::vector<double> h(2 * kTile * kTile);
    for (int a0 = 0; a0 < qx; a0 += kTile)
      for (int b0 = 0; b0 < qy; b0 += kTile) {
        hipLaunchKernelGGL((k_gram<kTile>), dim3(nparts_), dim3(256), 0,
                           c_->stream, n_, kTile, X + (size_t)a0 * n_, n_,
                           Y + (size_t)b0 * n_, n_, part_);
        hipLaunchKernelGGL(k_sum_cols, dim3(4), dim3(256), 0, c_->stream,
                           nparts_, 2 * kTile * kTile, part_, sum_);
        NH_HIP(hipGetLastError());
        NH_HIP(hipMemcpyAsync(h.data(), sum_, h.size() * sizeof(double),
                              hipMemcpyDeviceToHost, c_->stream));
        NH_HIP(hipStreamSynchronize(c_->stream));
        for (int a = 0; a < kTile && a0 + a < qx; ++a)
          for (int b = 0; b < kTile && b0 + b < qy; ++b)
            G[(size_t)(a0 + a) * qy + b0 + b] =
                cd(h[2 * (a * kTile + b)], h[2 * (a * kTile + b) + 1]);
      }
    return NH_OK;
  }

  // Y[:, 0..qy) = beta Y + X M with M (qx x qy), qy <= 16; Y must not
  // overlap X.
  nh_status gemm(const double2 *X, int qx, double2 *Y, int qy,
                 const std::vector<cd> &M, double beta) {
    NH_CHECK(qy >= 1 && qy <= 16, NH_ERR_ARG, "block update width %d", qy);
    int64_t g = (n_ + 255) / 256;
    if (g > 16384) g = 16384;
    for (int x0 = 0; x0 < qx; x0 += 16) {
      const int q = std::min(16, qx - x0);
      std::vector<double> h(2 * q * qy);
      for (int a = 0; a < q; ++a)
        for (int b = 0; b < qy; ++b) {
          h[2 * (a * qy + b)] = M[(size_t)(x0 + a) * qy + b].real();
          h[2 * (a * qy + b) + 1] = M[(size_t)(x0 + a) * qy + b].imag();
        }
      NH_HIP(hipMemcpyAsync(dM_, h.data(), h.size() * sizeof(double),
                            hipMemcpyHostToDevice, c_->stream));
      hipLaunchKernelGGL((k_block_gemm<16>), dim3((int)g), dim3(256), 0,
                         c_->stream, n_, q, qy, X + (size_t)x0 * n_, n_, Y, n_,
                         dM_, x0 == 0 ? beta : 1.0);
      NH_HIP(hipGetLastError());
      NH_HIP(hipStreamSynchronize(c_->stream));  // h is released on return
    }
    return NH_OK;
  }

  // Q R = W (k columns, k <= max_cols) by Cholesky QR applied twice; R =
  // R2 R1.  ok = false when W is numerically rank deficient (Q may then
  // hold partial results).  W is left unchanged, Q must not overlap W; both
  // have round_up(k, 8) columns allocated, the ones beyond k zero.
  nh_status cholqr(const double2 *W, double2 *Q, int k, std::vector<cd> &R,
                   bool &ok) {
    NH_CHECK(k >= 1 && k <= cap_, NH_ERR_ARG, "cholqr width %d", k);
    const int kp = (k + kTile - 1) / kTile * kTile;
    std::vector<cd> G, F[2];
    ok = true;
    for (int pass = 0; pass < 2; ++pass) {
      const double2 *src = pass == 0 ? W : Q;
      double2 *dst = pass == 0 ? Q : S_;
      NH_TRY(gram(src, kp, src, kp, G));
      F[pass].assign((size_t)k * k, 0.0);
      for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) F[pass][a * k + b] = G[(size_t)a * kp + b];
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
  int cap_;
  int nparts_ = 0;
  std::vector<void *> mem_;
  double *part_ = nullptr, *sum_ = nullptr, *dM_ = nullptr;
  double2 *S_ = nullptr;
};

}  // namespace nh
