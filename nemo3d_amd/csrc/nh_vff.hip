// Keating valence-force-field relaxation on the GPU.
//
// Energy (calc_estrain, NEMO_3D/nemo3d/src/base/strain3d.c:1739-1890), per
// atom i over its present neighbours n, p:
//   E_i = sum_n  a_in / (2 d0_in^2) (|r_in|^2 - d0_in^2)^2
//       + sum_{n<p} sqrt(b_in b_ip) / (|d0_in| |d0_ip|) (r_in.r_ip - d0_in.d0_ip)^2
// with r_in the strained bond vector, d0_in the ideal bond vector of the
// pair's material (unstrained lattice constant), a, b the Keating constants
// of the pair (StrainParam).  The total is sum_i E_i, so each bond term
// enters from both of its atoms.
//
// The gradient on atom i is gathered per atom (no atomics): its own bond
// and angle terms plus the angle terms centred on each neighbour j in which
// i is a leg.  Atoms with fewer than four neighbours are held fixed
// (Keating = FixSurfaceAll).  Minimisation: Polak-Ribiere+ nonlinear
// conjugate gradients with a secant line search on the directional
// derivative (the reference uses MacKay's macopt CG, eigen/src/macopt.c).
#include <chrono>
#include <cmath>
#include <vector>

#include "nh_internal.h"
#include "nh_vector.h"

namespace nh {

struct VffArgs {
  int A, S;
  const int32_t *species, *nbr, *bond_q;
  const double *alpha, *beta, *a0;  // [S][S]
  double lx, ly, lz;
};

__device__ __forceinline__ void bond_vec(const VffArgs &g, const double *u,
                                         int i, int n, int j, double r[3],
                                         double d0[3]) {
  const double lat[3] = {g.lx, g.ly, g.lz};
  const double a0 = g.a0[g.species[i] * g.S + g.species[j]];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double q = 0.25 * g.bond_q[(i * 4 + n) * 3 + k];
    r[k] = lat[k] * q + u[j * 3 + k] - u[i * 3 + k];
    d0[k] = a0 * q;
  }
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Gradient of E with respect to u_i (and E_i for the energy sum).
__global__ __launch_bounds__(kRedThreads) void k_vff_grad(
    VffArgs g, const double *__restrict__ u, double *__restrict__ grad,
    double *__restrict__ part) {
  const int i = blockIdx.x * kRedThreads + threadIdx.x;
  double e_i = 0.0, zero = 0.0;
  if (i < g.A) {
    double gr[3] = {0.0, 0.0, 0.0};
    double r[4][3], d0[4][3];
    int nb[4];
    int present = 0;
    const int ti = g.species[i];
    for (int n = 0; n < 4; ++n) {
      nb[n] = g.nbr[i * 4 + n];
      if (nb[n] >= 0) {
        bond_vec(g, u, i, n, nb[n], r[n], d0[n]);
        ++present;
      }
    }
    // terms centred on i
    for (int n = 0; n < 4; ++n) {
      if (nb[n] < 0) continue;
      const double al = g.alpha[ti * g.S + g.species[nb[n]]];
      const double d0sq = dot3(d0[n], d0[n]);
      const double dl = dot3(r[n], r[n]) - d0sq;
      e_i += 0.5 * al / d0sq * dl * dl;
      // this term and the same bond's term on the neighbour (equal value):
      // d/du_i of 2 * (a/2d0^2) dl^2 = -(4 a/d0^2) dl r_in
      const double f = 4.0 * al / d0sq * dl;
      for (int k = 0; k < 3; ++k) gr[k] -= f * r[n][k];
      for (int p = n + 1; p < 4; ++p) {
        if (nb[p] < 0) continue;
        const double be = sqrt(g.beta[ti * g.S + g.species[nb[n]]] *
                               g.beta[ti * g.S + g.species[nb[p]]]);
        const double K = be / sqrt(dot3(d0[p], d0[p]) * d0sq);
        const double t = dot3(r[n], r[p]) - dot3(d0[n], d0[p]);
        e_i += K * t * t;
        for (int k = 0; k < 3; ++k) gr[k] -= 2.0 * K * t * (r[n][k] + r[p][k]);
      }
    }
    // angle terms centred on each neighbour j with i as one leg
    for (int n = 0; n < 4; ++n) {
      const int j = nb[n];
      if (j < 0) continue;
      const int tj = g.species[j];
      // slot of i in j's list and the bond vector j -> i
      double rji[3], d0ji[3];
      int si = -1;
      for (int q = 0; q < 4; ++q)
        if (g.nbr[j * 4 + q] == i) si = q;
      bond_vec(g, u, j, si, i, rji, d0ji);
      const double b_ji = g.beta[tj * g.S + ti];
      for (int q = 0; q < 4; ++q) {
        const int k2 = g.nbr[j * 4 + q];
        if (q == si || k2 < 0) continue;
        double rjk[3], d0jk[3];
        bond_vec(g, u, j, q, k2, rjk, d0jk);
        const double be = sqrt(b_ji * g.beta[tj * g.S + g.species[k2]]);
        const double K = be / sqrt(dot3(d0ji, d0ji) * dot3(d0jk, d0jk));
        const double t = dot3(rji, rjk) - dot3(d0ji, d0jk);
        for (int k = 0; k < 3; ++k) gr[k] += 2.0 * K * t * rjk[k];
      }
    }
    const bool fixed = present < 4;
    for (int k = 0; k < 3; ++k) grad[i * 3 + k] = fixed ? 0.0 : gr[k];
  }
  if (part) {
    block_sum2(e_i, zero);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = e_i;
  }
}

namespace {

// sum_k a[k] b[k] partials over n doubles
__global__ __launch_bounds__(kRedThreads) void k_dotr(
    int64_t n, const double *__restrict__ a, const double *__restrict__ b,
    double *__restrict__ part) {
  double s = 0.0, z = 0.0;
  for (int64_t k = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kRedThreads)
    s += a[k] * b[k];
  block_sum2(s, z);
  if (threadIdx.x == 0) part[2 * blockIdx.x] = s;
}

// y = x + t d
__global__ void k_axpy_r(int64_t n, const double *__restrict__ x, double t,
                         const double *__restrict__ d, double *__restrict__ y) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    y[k] = x[k] + t * d[k];
}

// d = -g + bet d; bet == 0 does not read d (a fresh direction may sit in
// memory holding anything, and 0 * NaN is NaN)
__global__ void k_cg_dir(int64_t n, const double *__restrict__ g, double bet,
                         double *__restrict__ d) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    d[k] = bet == 0.0 ? -g[k] : -g[k] + bet * d[k];
}

class Vff {
 public:
  Vff(nh_context *c, VffArgs a) : c_(c), g_(a) {}
  ~Vff() {
    for (void *p : mem_) hipFree(p);
  }
  nh_status run(int max_iter, double gtol, double *out_disp, double *info);

 private:
  nh_status alloc(double **p, size_t n) {
    NH_TRY(device_malloc(c_, (void **)p, n * sizeof(double)));
    mem_.push_back(*p);
    NH_HIP(hipMemsetAsync(*p, 0, n * sizeof(double), c_->stream));
    return NH_OK;
  }
  nh_status grad(const double *u, double *gr, double *energy) {
    const int blocks = (g_.A + kRedThreads - 1) / kRedThreads;
    hipLaunchKernelGGL(k_vff_grad, dim3(blocks), dim3(kRedThreads), 0,
                       c_->stream, g_, u, gr, energy ? part_ : nullptr);
    NH_HIP(hipGetLastError());
    if (energy) NH_TRY(finish(blocks, energy));
    return NH_OK;
  }
  nh_status dot(const double *a, const double *b, double *out) {
    const int nb = reduce_blocks(n3_);
    hipLaunchKernelGGL(k_dotr, dim3(nb), dim3(kRedThreads), 0, c_->stream, n3_,
                       a, b, part_);
    NH_HIP(hipGetLastError());
    return finish(nb, out);
  }
  nh_status finish(int nparts, double *out) {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0,
                       c_->stream, nparts, part_, part_ + 2 * nparts_);
    NH_HIP(hipMemcpyAsync(c_->hred, part_ + 2 * nparts_, 2 * sizeof(double),
                          hipMemcpyDeviceToHost, c_->stream));
    NH_HIP(hipStreamSynchronize(c_->stream));
    *out = c_->hred[0];
    return NH_OK;
  }
  int grid(int64_t n) const {
    int64_t b = (n + 255) / 256;
    return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
  }

  nh_context *c_;
  VffArgs g_;
  std::vector<void *> mem_;
  int64_t n3_ = 0;
  int nparts_ = 0;
  double *part_ = nullptr;
};

nh_status Vff::run(int max_iter, double gtol, double *out_disp, double *info) {
  n3_ = (int64_t)g_.A * 3;
  nparts_ = std::max((g_.A + kRedThreads - 1) / kRedThreads, kRedMaxBlocks) + 8;
  NH_TRY(ensure_reduce(c_, 1));
  double *u, *ut, *gr, *gt, *d;
  NH_TRY(alloc(&u, n3_));
  NH_TRY(alloc(&ut, n3_));
  NH_TRY(alloc(&gr, n3_));
  NH_TRY(alloc(&gt, n3_));
  NH_TRY(alloc(&d, n3_));
  NH_TRY(alloc(&part_, (size_t)nparts_ * 2 + 16));
  NH_HIP(hipMemcpyAsync(u, c_->disp, n3_ * sizeof(double),
                        hipMemcpyDeviceToDevice, c_->stream));
  double e0 = 0.0, e = 0.0, gg = 0.0;
  NH_TRY(grad(u, gr, &e0));
  e = e0;
  NH_TRY(dot(gr, gr, &gg));
  const double g0 = std::sqrt(gg);
  hipLaunchKernelGGL(k_cg_dir, dim3(grid(n3_)), dim3(256), 0, c_->stream, n3_,
                     gr, 0.0, d);
  double step = 1e-4;
  int it = 0;
  for (; it < max_iter && std::sqrt(gg) > gtol; ++it) {
    double dg0;
    NH_TRY(dot(d, gr, &dg0));
    if (dg0 >= 0.0) {  // not a descent direction: restart along -g
      hipLaunchKernelGGL(k_cg_dir, dim3(grid(n3_)), dim3(256), 0, c_->stream,
                         n3_, gr, 0.0, d);
      NH_TRY(dot(d, gr, &dg0));
    }
    // secant search for phi'(t) = g(u + t d) . d = 0
    double t0 = 0.0, f0 = dg0, t1 = step, f1 = 0.0;
    for (int ls = 0; ls < 8; ++ls) {
      hipLaunchKernelGGL(k_axpy_r, dim3(grid(n3_)), dim3(256), 0, c_->stream,
                         n3_, u, t1, d, ut);
      NH_TRY(grad(ut, gt, nullptr));
      NH_TRY(dot(gt, d, &f1));
      if (std::fabs(f1) <= 1e-3 * std::fabs(dg0)) break;
      double tn;
      if (f1 < 0.0 && f1 <= f0) tn = 2.0 * t1;  // still descending: expand
      else tn = t1 - f1 * (t1 - t0) / (f1 - f0);
      if (!(tn > 0.0) || !std::isfinite(tn)) tn = 0.5 * t1;
      t0 = t1;
      f0 = f1;
      t1 = tn;
    }
    step = t1;
    hipLaunchKernelGGL(k_axpy_r, dim3(grid(n3_)), dim3(256), 0, c_->stream,
                       n3_, u, t1, d, u);
    // new gradient; Polak-Ribiere+
    NH_TRY(grad(u, gt, nullptr));
    double gg_new, g_old_new;
    NH_TRY(dot(gt, gt, &gg_new));
    NH_TRY(dot(gt, gr, &g_old_new));
    double bet = (gg_new - g_old_new) / gg;
    if (bet < 0.0) bet = 0.0;
    std::swap(gr, gt);
    gg = gg_new;
    hipLaunchKernelGGL(k_cg_dir, dim3(grid(n3_)), dim3(256), 0, c_->stream,
                       n3_, gr, bet, d);
    NH_HIP(hipGetLastError());
  }
  NH_TRY(grad(u, gt, &e));
  NH_HIP(hipMemcpyAsync(c_->disp, u, n3_ * sizeof(double),
                        hipMemcpyDeviceToDevice, c_->stream));
  if (out_disp)
    NH_HIP(hipMemcpyAsync(out_disp, u, n3_ * sizeof(double),
                          hipMemcpyDeviceToHost, c_->stream));
  NH_HIP(hipStreamSynchronize(c_->stream));
  info[0] = it;
  info[1] = e0;
  info[2] = e;
  info[3] = g0;
  info[4] = std::sqrt(gg);
  return NH_OK;
}

}  // namespace
}  // namespace nh

extern "C" nh_status nh_vff_energy_grad(nh_context *ctx, const double *alpha,
                                        const double *beta, const double *a0,
                                        const double *disp, double *grad_out,
                                        double *energy) {
  NH_CHECK(ctx && alpha && beta && a0 && disp && grad_out && energy,
           NH_ERR_ARG, "nh_vff_energy_grad: bad arguments");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  const int S = ctx->n_species, A = ctx->n_atoms;
  NH_TRY(nh::ensure_reduce(ctx, 1));
  std::vector<double *> m(6, nullptr);
  const size_t sz[6] = {(size_t)S * S, (size_t)S * S, (size_t)S * S,
                        (size_t)A * 3, (size_t)A * 3,
                        (size_t)(A / nh::kRedThreads + 2) * 2 + 8};
  for (int k = 0; k < 6; ++k) NH_TRY(nh::device_malloc(ctx, (void **)&m[k], sz[k] * sizeof(double)));
  hipMemcpy(m[0], alpha, sz[0] * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(m[1], beta, sz[1] * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(m[2], a0, sz[2] * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(m[3], disp, sz[3] * sizeof(double), hipMemcpyHostToDevice);
  nh::VffArgs a{A, S, ctx->species, ctx->nbr, ctx->bond_q, m[0], m[1], m[2],
                ctx->lattice[0], ctx->lattice[1], ctx->lattice[2]};
  const int blocks = (A + nh::kRedThreads - 1) / nh::kRedThreads;
  hipLaunchKernelGGL(nh::k_vff_grad, dim3(blocks), dim3(nh::kRedThreads), 0,
                     ctx->stream, a, m[3], m[4], m[5]);
  hipLaunchKernelGGL(nh::k_sum_partials, dim3(1), dim3(nh::kRedThreads), 0,
                     ctx->stream, blocks, m[5], m[5] + 2 * (blocks + 1));
  hipError_t e = hipGetLastError();
  double res[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpy(grad_out, m[4], sz[4] * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(res, m[5] + 2 * (blocks + 1), 2 * sizeof(double), hipMemcpyDeviceToHost);
  for (double *p : m) hipFree(p);
  NH_CHECK(e == hipSuccess, NH_ERR_HIP, "nh_vff_energy_grad: %s", hipGetErrorString(e));
  *energy = res[0];
  return NH_OK;
}

extern "C" nh_status nh_vff_relax(nh_context *ctx, const double *alpha,
                                  const double *beta, const double *a0,
                                  int max_iter, double gtol, double *disp_out,
                                  double *info) {
  NH_CHECK(ctx && alpha && beta && a0 && info && max_iter >= 0 && gtol >= 0,
           NH_ERR_ARG, "nh_vff_relax: bad arguments");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  const int S = ctx->n_species;
  double *dal = nullptr, *dbe = nullptr, *da0 = nullptr;
  NH_TRY(nh::device_malloc(ctx, (void **)&dal, S * S * sizeof(double)));
  NH_TRY(nh::device_malloc(ctx, (void **)&dbe, S * S * sizeof(double)));
  NH_TRY(nh::device_malloc(ctx, (void **)&da0, S * S * sizeof(double)));
  hipMemcpy(dal, alpha, S * S * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(dbe, beta, S * S * sizeof(double), hipMemcpyHostToDevice);
  hipMemcpy(da0, a0, S * S * sizeof(double), hipMemcpyHostToDevice);
  nh::VffArgs a{ctx->n_atoms, S, ctx->species, ctx->nbr, ctx->bond_q,
                dal, dbe, da0, ctx->lattice[0], ctx->lattice[1],
                ctx->lattice[2]};
  nh_status st;
  {
    nh::Vff v(ctx, a);
    st = v.run(max_iter, gtol, disp_out, info);
  }
  hipFree(dal);
  hipFree(dbe);
  hipFree(da0);
  return st;
}
