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
// The gradient is the reference's (strain3d.c:1806-1830 and 1889-2050): an
// atom receives the derivative of its own copy of each bond term only
// (-4 a/(2 d0^2) dl r_in), half the derivative of E, and the full
// derivative of the angle terms.  The relaxation follows that gradient, so
// its stationary point is that of E with each bond term counted once; E
// itself is reported as the reference prints it.
//
// StrainModel = VFF_keating_strained replaces a and sqrt(b_in b_ip) by the
// strain-scaled constants of StrainParam::alpha_eff / beta_eff
// (StrainParam.c:60-214): per bond a (1 - A (|r|^2 / d0^2 - 1)), per angle
// b_in (1 - B (cos + 1/3)) (1 - C (|r_in| |r_ip| / (|d0_in| |d0_ip|) - 1))
// with B, C the geometric means of the two bonds' constants and both
// factors truncated near their extrema; the reference's b_in is the
// constant of the lower-slot leg n.  As in calc_estrain (strain3d.c:
// 1687-1700, 1835-1860) the scaled constants are taken as fixed when the
// gradient is formed, so the "gradient" of this model is that of the
// harmonic form at the current constants.
//
// The gradient on atom i is gathered per atom (no atomics): its own bond
// and angle terms plus the angle terms centred on each neighbour j in which
// i is a leg.  Keating = FixSurfaceAll holds atoms with fewer than four
// neighbours fixed; AllFree moves every atom.  Minimisation: Polak-Ribiere+
// nonlinear conjugate gradients with a secant line search on the
// directional derivative (the reference uses MacKay's macopt CG,
// eigen/src/macopt.c, whose line search also uses only gradients).
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
  const double *aa, *ba, *ca;  // [S][S], NH_VFF_KEATING_STRAINED only
  int model;
  int fix_surface;
};

// StrainParam::alpha_eff with anh = 1 (StrainParam.c:60-100); the
// reference's how_close_to_maximum_of_distorsion_energy is 0.001
constexpr double kVffNearMax = 0.001;

__device__ __forceinline__ double alpha_eff(double al0, double A, double r2,
                                            double d2) {
  const double delt = 1.0 - kVffNearMax;
  if (r2 < d2 * (1.0 + delt * 2.0 / (3.0 * A)))
    return al0 * (1.0 - A * (r2 / d2 - 1.0));
  return al0 * (1.0 - delt / 1.5);
}

// StrainParam::beta_eff with anh = 1 (StrainParam.c:118-214): be0 the
// constant of the first leg, d1/d2 the ideal and r1/r2 the strained legs
__device__ __forceinline__ double beta_eff(double be0, double B, double C,
                                           const double d1[3],
                                           const double d2[3],
                                           const double r1[3],
                                           const double r2[3]) {
  const double d10 = sqrt(d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2]);
  const double d20 = sqrt(d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2]);
  const double rr = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  const double r10 = sqrt(r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2]);
  const double r20 = sqrt(r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2]);
  const double rd = r10 * r20 / (d10 * d20);
  const double tt = rr / (r10 * r20);
  const double delt = 1.0 + kVffNearMax, delh = 1.0 - kVffNearMax;
  double beta = be0;
  bool flag = true;
  if (tt > delt / 3.0 * (2.0 * (1.0 / B - 1.0 / 3.0) - 1.0 / (3.0 * rd))) {
    beta = be0 * ((1.0 - 2.0 * delt / 3.0) +
                  B / 3.0 * (delt / 3.0 * (1.0 / rd + 2.0) - 1.0)) *
           (1.0 - C * (rd - 1.0));
    flag = false;
  }
  if (rd > 1.0 + delh * 2.0 / (3.0 * C)) {
    const double be = be0 * (1.0 - B * (tt + 1.0 / 3.0)) * (1.0 - delh * 2.0 / 3.0);
    if (flag || be > beta) beta = be;
    flag = false;
  }
  if (flag) beta = be0 * (1.0 - B * (tt + 1.0 / 3.0)) * (1.0 - C * (rd - 1.0));
  return beta;
}

// angle constant of the legs n1 < n2 (slots) centred on an atom of type tc
__device__ __forceinline__ double angle_const(const VffArgs &g, int tc, int t1,
                                              int t2, const double d1[3],
                                              const double d2[3],
                                              const double r1[3],
                                              const double r2[3]) {
  const double b1 = g.beta[tc * g.S + t1], b2 = g.beta[tc * g.S + t2];
  if (g.model == 0) return sqrt(b1 * b2);
  const double B = sqrt(g.ba[tc * g.S + t1] * g.ba[tc * g.S + t2]);
  const double C = sqrt(g.ca[tc * g.S + t1] * g.ca[tc * g.S + t2]);
  return beta_eff(sqrt(b1 * b1), B, C, d1, d2, r1, r2);
}

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
      const int tn = g.species[nb[n]];
      const double d0sq = dot3(d0[n], d0[n]);
      const double rsq = dot3(r[n], r[n]);
      const double al =
          g.model == 0 ? g.alpha[ti * g.S + tn]
                       : alpha_eff(g.alpha[ti * g.S + tn], g.aa[ti * g.S + tn],
                                   rsq, d0sq);
      const double dl = rsq - d0sq;
      e_i += 0.5 * al / d0sq * dl * dl;
      // the derivative of this atom's copy of the bond term: d/du_i of
      // (a/2d0^2) dl^2 = -(2 a/d0^2) dl r_in (the neighbour's copy is
      // not differentiated here, as in the reference)
      const double f = 2.0 * al / d0sq * dl;
      for (int k = 0; k < 3; ++k) gr[k] -= f * r[n][k];
      for (int p = n + 1; p < 4; ++p) {
        if (nb[p] < 0) continue;
        const double be = angle_const(g, ti, tn, g.species[nb[p]], d0[n],
                                      d0[p], r[n], r[p]);
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
      for (int q = 0; q < 4; ++q) {
        const int k2 = g.nbr[j * 4 + q];
        if (q == si || k2 < 0) continue;
        double rjk[3], d0jk[3];
        bond_vec(g, u, j, q, k2, rjk, d0jk);
        const int tk = g.species[k2];
        // the angle's legs in j's slot order
        const double be =
            si < q ? angle_const(g, tj, ti, tk, d0ji, d0jk, rji, rjk)
                   : angle_const(g, tj, tk, ti, d0jk, d0ji, rjk, rji);
        const double K = be / sqrt(dot3(d0ji, d0ji) * dot3(d0jk, d0jk));
        const double t = dot3(rji, rjk) - dot3(d0ji, d0jk);
        for (int k = 0; k < 3; ++k) gr[k] += 2.0 * K * t * rjk[k];
      }
    }
    const bool fixed = g.fix_surface && present < 4;
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
    for (void *p : mem_) (void)hipFree(p);
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

namespace nh {
namespace {

// The parameter tables of a model on the device (six [S][S] tables in one
// allocation), freed with the object.
class VffTables {
 public:
  ~VffTables() { (void)hipFree(mem_); }
  nh_status upload(nh_context *c, const nh_vff_params *p) {
    NH_CHECK(p && p->alpha && p->beta && p->a0 &&
                 (p->model == NH_VFF_KEATING ||
                  (p->model == NH_VFF_KEATING_STRAINED && p->aa && p->ba && p->ca)) &&
                 (p->fix_surface == 0 || p->fix_surface == 1),
             NH_ERR_ARG, "VFF: bad parameters (model %d, fix_surface %d)",
             p ? p->model : -1, p ? p->fix_surface : -1);
    const size_t ss = (size_t)c->n_species * c->n_species;
    NH_TRY(device_malloc(c, (void **)&mem_, 6 * ss * sizeof(double)));
    const double *src[6] = {p->alpha, p->beta, p->a0, p->aa, p->ba, p->ca};
    for (int k = 0; k < 6; ++k)
      if (src[k])
        NH_HIP(hipMemcpy(mem_ + k * ss, src[k], ss * sizeof(double),
                         hipMemcpyHostToDevice));
    args = VffArgs{c->n_atoms, c->n_species, c->species, c->nbr, c->bond_q,
                   mem_, mem_ + ss, mem_ + 2 * ss, c->lattice[0], c->lattice[1],
                   c->lattice[2], mem_ + 3 * ss, mem_ + 4 * ss, mem_ + 5 * ss,
                   p->model, p->fix_surface};
    return NH_OK;
  }
  VffArgs args{};

 private:
  double *mem_ = nullptr;
};

nh_status energy_grad(nh_context *ctx, const nh_vff_params *p,
                      const double *disp, double *grad_out, double *energy) {
  NH_CHECK(ctx && disp && grad_out && energy, NH_ERR_ARG,
           "nh_vff_energy_grad: bad arguments");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  VffTables t;
  NH_TRY(t.upload(ctx, p));
  const int A = ctx->n_atoms;
  const int blocks = (A + kRedThreads - 1) / kRedThreads;
  double *m[3] = {nullptr, nullptr, nullptr};
  const size_t sz[3] = {(size_t)A * 3, (size_t)A * 3, (size_t)(blocks + 2) * 2 + 8};
  nh_status st = NH_OK;
  for (int k = 0; k < 3 && st == NH_OK; ++k)
    st = device_malloc(ctx, (void **)&m[k], sz[k] * sizeof(double));
  hipError_t e = hipSuccess;
  double res[2] = {0, 0};
  if (st == NH_OK) {
    e = hipMemcpy(m[0], disp, sz[0] * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_vff_grad, dim3(blocks), dim3(kRedThreads), 0,
                         ctx->stream, t.args, m[0], m[1], m[2]);
      hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedThreads), 0,
                         ctx->stream, blocks, m[2], m[2] + 2 * (blocks + 1));
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpy(grad_out, m[1], sz[1] * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(res, m[2] + 2 * (blocks + 1), 2 * sizeof(double),
                    hipMemcpyDeviceToHost);
  }
  for (double *q : m) (void)hipFree(q);
  NH_TRY(st);
  NH_CHECK(e == hipSuccess, NH_ERR_HIP, "nh_vff_energy_grad: %s", hipGetErrorString(e));
  *energy = res[0];
  return NH_OK;
}

nh_status relax(nh_context *ctx, const nh_vff_params *p, int max_iter,
                double gtol, double *disp_out, double *info) {
  NH_CHECK(ctx && info && max_iter >= 0 && gtol >= 0, NH_ERR_ARG,
           "nh_vff_relax: bad arguments");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  VffTables t;
  NH_TRY(t.upload(ctx, p));
  Vff v(ctx, t.args);
  return v.run(max_iter, gtol, disp_out, info);
}

nh_vff_params keating(const double *alpha, const double *beta, const double *a0) {
  return nh_vff_params{alpha, beta, a0, nullptr, nullptr, nullptr,
                       NH_VFF_KEATING, 1};
}

}  // namespace
}  // namespace nh

extern "C" nh_status nh_vff_energy_grad(nh_context *ctx, const double *alpha,
                                        const double *beta, const double *a0,
                                        const double *disp, double *grad_out,
                                        double *energy) {
  const nh_vff_params p = nh::keating(alpha, beta, a0);
  return nh::energy_grad(ctx, &p, disp, grad_out, energy);
}

extern "C" nh_status nh_vff_relax(nh_context *ctx, const double *alpha,
                                  const double *beta, const double *a0,
                                  int max_iter, double gtol, double *disp_out,
                                  double *info) {
  const nh_vff_params p = nh::keating(alpha, beta, a0);
  return nh::relax(ctx, &p, max_iter, gtol, disp_out, info);
}

extern "C" nh_status nh_vff_energy_grad_model(nh_context *ctx,
                                              const nh_vff_params *p,
                                              const double *disp,
                                              double *grad_out, double *energy) {
  return nh::energy_grad(ctx, p, disp, grad_out, energy);
}

extern "C" nh_status nh_vff_relax_model(nh_context *ctx, const nh_vff_params *p,
                                        int max_iter, double gtol,
                                        double *disp_out, double *info) {
  return nh::relax(ctx, p, max_iter, gtol, disp_out, info);
}
