// Operator callbacks with the reference's signatures, so that its unchanged
// solver drivers (eigensolve_lanczos, eigensolve_blk_lanczos, the TRACEMIN
// reverse-communication loop) run on the device operator.  The argList
// element 0 (the reference's qd_struct) and the `d` argument cannot carry a
// libnemohip context; the context is bound once with nh_shim_bind.  Like the
// functions they replace, these return nothing and abort with a message on
// error (die(), h_cvectr_mult.c:312-321).
#include <cstdlib>
#include <vector>

#include "nh_internal.h"

namespace nh {
namespace {

struct Shim {
  nh_context *ctx = nullptr;
  // registered wave functions (the reference's d->WF table) for the spectral
  // projection of matvect_mult
  nh_cplx *wf = nullptr;  // device [nwf][n]
  std::vector<double> energy;
  int64_t n = 0;
};

Shim g_shim;

[[noreturn]] void die(const char *where) {
  std::fprintf(stderr, "libnemohip %s: %s\n", where, nh_last_error());
  std::fflush(stderr);
  std::abort();
}

// Y_j = (H - shift) X_j for row-pointer matrices (cmatrix: one row per
// vector), then y -= sum_k E_k <psi_k|x> psi_k over the listed projectors.
nh_status apply(double shift, nh_cplx *const *Y, const nh_cplx *const *X,
                int ncol, int nproj, const int *which) {
  nh_context *c = g_shim.ctx;
  NH_CHECK(c, NH_ERR_STATE, "no context bound (nh_shim_bind)");
  NH_CHECK(ncol >= 1 && Y && X, NH_ERR_ARG, "bad vectors");
  NH_CHECK(nproj >= 0, NH_ERR_ARG, "negative Nproj");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(c->device));
  const int64_t n = c->n;
  NH_TRY(ensure_buffers(c, n * ncol));
  for (int j = 0; j < ncol; ++j)
    NH_HIP(hipMemcpyAsync(c->xbuf + j * n, X[j], n * sizeof(nh_cplx),
                          hipMemcpyHostToDevice, c->stream));
  NH_TRY(launch_hv(c, shift, c->xbuf, c->ybuf, ncol, n));
  if (nproj > 0) {
    NH_CHECK(ncol == 1 && which, NH_ERR_ARG, "projection needs one vector and "
             "the eval_scoperti list");
    NH_CHECK(g_shim.wf && g_shim.n == n, NH_ERR_STATE,
             "projection requested but no wave functions registered");
    for (int k = 0; k < nproj; ++k) {
      const int w = which[k];
      NH_CHECK(w >= 0 && w < (int)g_shim.energy.size(), NH_ERR_ARG,
               "eval_scoperti[%d] = %d out of range", k, w);
      double ip[2];
      const nh_cplx *psi = g_shim.wf + (int64_t)w * n;
      NH_TRY(reduce_dot(c, n, psi, c->xbuf, ip));
      const nh_cplx a = {-g_shim.energy[w] * ip[0], -g_shim.energy[w] * ip[1]};
      NH_TRY(launch_axpby(c, n, a, psi, nh_cplx{1.0, 0.0}, c->ybuf));
    }
  }
  for (int j = 0; j < ncol; ++j)
    NH_HIP(hipMemcpyAsync(Y[j], c->ybuf + j * n, n * sizeof(nh_cplx),
                          hipMemcpyDeviceToHost, c->stream));
  NH_HIP(hipStreamSynchronize(c->stream));
  return NH_OK;
}

}  // namespace
}  // namespace nh

extern "C" {

nh_status nh_shim_bind(nh_context *ctx) {
  using nh::g_shim;
  if (g_shim.wf) {
    (void)hipFree(g_shim.wf);
    g_shim.wf = nullptr;
  }
  g_shim.energy.clear();
  g_shim.ctx = ctx;
  g_shim.n = ctx ? ctx->n : 0;
  return NH_OK;
}

nh_status nh_shim_set_wavefunctions(int nwf, const double *energy,
                                    const nh_cplx *psi) {
  using nh::g_shim;
  nh_context *c = g_shim.ctx;
  NH_CHECK(c, NH_ERR_STATE, "no context bound (nh_shim_bind)");
  NH_CHECK(nwf >= 0 && (nwf == 0 || (energy && psi)), NH_ERR_ARG,
           "bad wave function table");
  NH_HIP(hipSetDevice(c->device));
  if (g_shim.wf) (void)hipFree(g_shim.wf);
  g_shim.wf = nullptr;
  g_shim.energy.assign(energy, energy + nwf);
  if (nwf == 0) return NH_OK;
  const size_t bytes = (size_t)nwf * c->n * sizeof(nh_cplx);
  NH_TRY(nh::device_malloc(c, (void **)&g_shim.wf, bytes));
  NH_HIP(hipMemcpy(g_shim.wf, psi, bytes, hipMemcpyHostToDevice));
  g_shim.n = c->n;
  return NH_OK;
}

void nh_matvect_mult(const int **argList, nh_cplx *y, const nh_cplx *x) {
  const double shift = *reinterpret_cast<const double *>(argList[1]);
  const int nproj = *reinterpret_cast<const int *>(argList[2]);
  const int *which =
      nproj > 0 ? *reinterpret_cast<int *const *>(argList[3]) : nullptr;
  nh_cplx *Y[1] = {y};
  const nh_cplx *X[1] = {x};
  if (nh::apply(shift, Y, X, 1, nproj, which) != NH_OK)
    nh::die("matvect_mult");
}

void nh_matmat_mult(const int **argList, nh_cplx **y, nh_cplx *const *x) {
  // argList[3] (Nproj) is not applied, as in the reference (h_cmat_mult.c:29)
  const double shift = *reinterpret_cast<const double *>(argList[1]);
  const int ncol = *reinterpret_cast<const int *>(argList[2]);
  if (nh::apply(shift, y, x, ncol, 0, nullptr) != NH_OK)
    nh::die("matmat_mult");
}

void nh_Hmult_spds_complete(nh_cplx *y, void *d, double shift, nh_cplx *x) {
  (void)d;
  nh_cplx *Y[1] = {y};
  const nh_cplx *X[1] = {x};
  if (nh::apply(shift, Y, X, 1, 0, nullptr) != NH_OK)
    nh::die("Hmult_spds_complete");
}

void nh_Hmatmult_spds_complete(nh_cplx **Y, void *d, double shift,
                               nh_cplx **X, int num_col) {
  (void)d;
  if (nh::apply(shift, Y, X, num_col, 0, nullptr) != NH_OK)
    nh::die("Hmatmult_spds_complete");
}

}  // extern "C"
