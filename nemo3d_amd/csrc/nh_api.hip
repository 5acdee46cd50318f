// C-ABI entry points of libnemohip (declared in include/nemohip.h).
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "nh_internal.h"

namespace nh {

static thread_local char g_err[1024] = "";

// Every live context, so that an allocation that fails on a device can take
// back the idle workspace blocks of all contexts on that device, and ring
// sizing can count them as available.
static std::mutex g_ctx_mu;
static std::vector<nh_context *> g_ctxs;

Progress::Progress() {
  const char *e = std::getenv("NEMOHIP_PROGRESS");
  on = e && e[0] == '1';
  last = std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch()).count();
}

void Progress::tick(const char *what, int iter, int extra) {
  if (!on) return;
  const double now = std::chrono::duration<double>(
                         std::chrono::steady_clock::now().time_since_epoch()).count();
  if (now - last < 15.0) return;
  last = now;
  std::fprintf(stderr, "# %s: iteration %d (%d)\n", what, iter, extra);
  std::fflush(stderr);
}

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// Frees the idle workspace blocks of one context (after its queued work).
static nh_status pool_trim(nh_context *c) {
  NH_HIP(hipStreamSynchronize(c->stream));
  std::vector<nh_context::PoolBlock> keep;
  for (auto &b : c->pool) {
    if (b.used) keep.push_back(b);
    else NH_HIP(hipFree(b.p));
  }
  c->pool.swap(keep);
  return NH_OK;
}

nh_status device_malloc(nh_context *c, void **p, size_t bytes) {
  bytes = bytes > 0 ? bytes : 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    // give the idle workspace of every context on this device back and
    // try once more
    (void)hipGetLastError();
    std::vector<nh_context *> same;
    {
      std::lock_guard<std::mutex> g(g_ctx_mu);
      for (nh_context *o : g_ctxs)
        if (o->device == c->device) same.push_back(o);
    }
    for (nh_context *o : same) NH_TRY(pool_trim(o));
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return NH_ERR_NOMEM;
  }
  return NH_OK;
}

nh_status pool_alloc(nh_context *c, void **p, size_t bytes) {
  bytes = bytes > 0 ? (bytes + 255) / 256 * 256 : 256;
  int best = -1;
  for (int i = 0; i < (int)c->pool.size(); ++i) {
    const auto &b = c->pool[i];
    if (!b.used && b.bytes >= bytes && b.bytes <= 2 * bytes &&
        (best < 0 || b.bytes < c->pool[best].bytes))
      best = i;
  }
  if (best >= 0) {
    c->pool[best].used = true;
    *p = c->pool[best].p;
    return NH_OK;
  }
  NH_TRY(device_malloc(c, p, bytes));
  c->pool.push_back({*p, bytes, true});
  return NH_OK;
}

void pool_release(nh_context *c, void *p) {
  for (auto &b : c->pool)
    if (b.p == p) {
      b.used = false;
      return;
    }
}

size_t pool_free_bytes(const nh_context *c) {
  std::lock_guard<std::mutex> g(g_ctx_mu);
  size_t s = 0;
  for (const nh_context *o : g_ctxs)
    if (o->device == c->device)
      for (const auto &b : o->pool)
        if (!b.used) s += b.bytes;
  return s;
}

nh_status ensure_buffers(nh_context *c, int64_t elems) {
  if (elems <= c->buf_cap) return NH_OK;
  NH_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(c->xbuf);
  (void)hipFree(c->ybuf);
  c->xbuf = c->ybuf = nullptr;
  c->buf_cap = 0;
  NH_TRY(device_malloc(c, (void **)&c->xbuf, elems * sizeof(nh_cplx)));
  NH_TRY(device_malloc(c, (void **)&c->ybuf, elems * sizeof(nh_cplx)));
  c->buf_cap = elems;
  return NH_OK;
}

template <typename T>
static nh_status upload(T **dst, const T *src, size_t count) {
  NH_HIP(hipMalloc(dst, count * sizeof(T) + 16));
  NH_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return NH_OK;
}

static void destroy(nh_context *c) {
  if (!c) return;
  {
    std::lock_guard<std::mutex> g(g_ctx_mu);
    g_ctxs.erase(std::remove(g_ctxs.begin(), g_ctxs.end(), c), g_ctxs.end());
  }
  (void)hipSetDevice(c->device);
  void *ptrs[] = {c->parmat, c->strnmat, c->species, c->is_cation, c->nbr,
                  c->bond_q, c->disp, c->phi, c->cations, c->anions, c->bond_of,
                  c->rev_slot, c->onsite, c->blocks64, c->blocks32,
                  c->onsite32, c->sk64, c->sk32, c->corr, c->red, c->xbuf,
                  c->ybuf};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  (void)hipStreamSynchronize(c->stream);
  for (auto &b : c->pool) (void)hipFree(b.p);
  if (c->hred) (void)hipHostFree(c->hred);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// Orders an atom list by 32x32x32-cell bricks (bricks in z, y, x order,
// cells and atoms in their original order inside a brick).  The SK kernels give
// consecutive workgroups consecutive list entries and each XCD a contiguous
// part of the list, so a brick's atoms and most of their neighbours are in
// flight together and the neighbours' vector entries are reused from the
// XCD's L2 instead of being fetched again when the next plane of cells is
// processed.  Ideal positions (units a/4) come from the ideal bond vectors
// by a walk over the bond graph.
static void brick_order(const nh_model *m, std::vector<int32_t> &list) {
  const int A = m->n_atoms;
  std::vector<int32_t> pos((size_t)A * 3, 0);
  std::vector<char> seen(A, 0);
  std::vector<int32_t> stack;
  for (int s0 = 0; s0 < A; ++s0) {
    if (seen[s0]) continue;
    seen[s0] = 1;
    stack.push_back(s0);
    while (!stack.empty()) {
      const int i = stack.back();
      stack.pop_back();
      for (int n = 0; n < 4; ++n) {
        const int j = m->nbr[i * 4 + n];
        if (j < 0 || seen[j]) continue;
        seen[j] = 1;
        for (int d = 0; d < 3; ++d)
          pos[(size_t)j * 3 + d] = pos[(size_t)i * 3 + d] + m->bond_q[(i * 4 + n) * 3 + d];
        stack.push_back(j);
      }
    }
  }
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
  for (int i = 0; i < A; ++i)
    for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], pos[(size_t)i * 3 + d]);
  // 32 cells of 4 units (a/4).  Measured on the 786,520-atom dot (FP32 /
  // FP64 fused Lanczos H*v): 4-cell bricks 0.241 / -, 8-cell 0.239 / 0.338,
  // 16-cell 0.234 / 0.337, 32-cell 0.229-0.232 / 0.337-0.338, 64-cell
  // 0.234 / 0.339 ms; octants at the medians (one per XCD) with 8-cell
  // bricks inside 0.245 / 0.343
  constexpr int kBrick = 128;
  auto key = [&](int32_t i) {
    int64_t b[3];
    for (int d = 0; d < 3; ++d) b[d] = (pos[(size_t)i * 3 + d] - lo[d]) / kBrick;
    return (b[2] << 42) | (b[1] << 21) | b[0];
  };
  std::stable_sort(list.begin(), list.end(),
                   [&](int32_t a, int32_t b) { return key(a) < key(b); });
}

static nh_status create(const nh_model *m, int device, nh_context **out) {
  NH_CHECK(m && out, NH_ERR_ARG, "null model or output pointer");
  NH_CHECK(m->nb == 10 || m->nb == 20, NH_ERR_ARG, "nb must be 10 or 20");
  NH_CHECK(m->n_atoms > 0 && m->n_species > 0, NH_ERR_ARG, "empty model");
  NH_CHECK(m->parmat && m->strnmat && m->species && m->is_cation && m->nbr &&
               m->bond_q && m->disp,
           NH_ERR_ARG, "missing model array");
  const int A = m->n_atoms, S = m->n_species;
  NH_CHECK((int64_t)A * m->nb < ((int64_t)1 << 40), NH_ERR_ARG, "too large");
  std::vector<int32_t> cations, anions, bond_of((size_t)A * 4, -1),
      rev((size_t)A * 4, -1);
  for (int i = 0; i < A; ++i) {
    NH_CHECK(m->species[i] >= 0 && m->species[i] < S, NH_ERR_ARG,
             "species[%d] out of range", i);
    int present = 0;
    for (int n = 0; n < 4; ++n) {
      const int j = m->nbr[i * 4 + n];
      NH_CHECK(j >= -1 && j < A && j != i, NH_ERR_ARG, "nbr[%d][%d] invalid",
               i, n);
      if (j < 0) continue;
      ++present;
      NH_CHECK((m->is_cation[i] != 0) != (m->is_cation[j] != 0), NH_ERR_ARG,
               "bond %d-%d does not join a cation and an anion", i, j);
      int r = -1;
      for (int q = 0; q < 4; ++q)
        if (m->nbr[j * 4 + q] == i) r = q;
      NH_CHECK(r >= 0, NH_ERR_ARG, "neighbour table not symmetric at %d-%d",
               i, j);
      rev[i * 4 + n] = r;
    }
    NH_CHECK(present > 0, NH_ERR_ARG, "atom %d has no neighbours", i);
  }
  for (int i = 0; i < A; ++i)
    (m->is_cation[i] ? cations : anions).push_back(i);
  brick_order(m, cations);
  brick_order(m, anions);
  for (size_t ci = 0; ci < cations.size(); ++ci) {
    const int i = cations[ci];
    for (int n = 0; n < 4; ++n) {
      const int j = m->nbr[i * 4 + n];
      if (j < 0) continue;
      bond_of[i * 4 + n] = (int32_t)(ci * 4 + n);
      bond_of[j * 4 + rev[i * 4 + n]] = (int32_t)(ci * 4 + n);
    }
  }
  NH_CHECK(!cations.empty(), NH_ERR_ARG, "no cation-role atoms");

  NH_HIP(hipSetDevice(device));
  NH_TRY(init_noalloc_scratch(device));
  nh_context *c = new nh_context();
  c->device = device;
  c->nb = m->nb;
  c->n_atoms = A;
  c->n_species = S;
  c->n_cations = (int)cations.size();
  c->n_anions = (int)anions.size();
  c->n = (int64_t)A * m->nb;
  for (int k = 0; k < 3; ++k) c->lattice[k] = m->lattice[k];
  c->passivate = m->passivate;
  c->db_shift = m->db_shift;
  nh_status st = NH_OK;
#define UP(dst, src, cnt)                              \
  if (st == NH_OK) st = upload(&c->dst, src, (size_t)(cnt));
  UP(parmat, m->parmat, (size_t)S * S * 35);
  UP(strnmat, m->strnmat, (size_t)S * S * 33);
  UP(species, m->species, A);
  UP(is_cation, m->is_cation, A);
  UP(nbr, m->nbr, (size_t)A * 4);
  UP(bond_q, m->bond_q, (size_t)A * 12);
  UP(disp, m->disp, (size_t)A * 3);
  if (m->phi) UP(phi, m->phi, A);
  UP(cations, cations.data(), cations.size());
  if (!anions.empty()) UP(anions, anions.data(), anions.size());
  UP(bond_of, bond_of.data(), (size_t)A * 4);
  UP(rev_slot, rev.data(), (size_t)A * 4);
#undef UP
  if (st == NH_OK) {
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
      set_error("hipStreamCreate: %s", hipGetErrorString(e));
      st = NH_ERR_HIP;
    }
  }
  if (st != NH_OK) {
    destroy(c);
    return st;
  }
  {
    std::lock_guard<std::mutex> g(g_ctx_mu);
    g_ctxs.push_back(c);
  }
  *out = c;
  return NH_OK;
}

static nh_status check_ctx(nh_context *c) {
  NH_CHECK(c, NH_ERR_ARG, "null context");
  (void)hipGetLastError();  // errors of earlier, already reported calls
  NH_HIP(hipSetDevice(c->device));
  return NH_OK;
}

}  // namespace nh

using namespace nh;

extern "C" {

nh_status nh_context_set(nh_context *ctx, int what, int value) {
  NH_CHECK(ctx, NH_ERR_ARG, "null context");
  switch (what) {
    case NH_SET_HV_STAGED:
      NH_CHECK(value >= -1 && value <= 1, NH_ERR_ARG,
               "NH_SET_HV_STAGED: value %d not in -1..1", value);
      ctx->hv_staged = value;
      return NH_OK;
    case NH_SET_SK_LANES:
      NH_CHECK(value == -1 || value == 4 || value == 8, NH_ERR_ARG,
               "NH_SET_SK_LANES: value %d not -1, 4 or 8", value);
      ctx->sk_lanes = value;
      return NH_OK;
    case NH_SET_LZ_VECTORS:
      NH_CHECK(value >= -1 && value <= 1, NH_ERR_ARG,
               "NH_SET_LZ_VECTORS: value %d not in -1..1", value);
      ctx->lz_vectors = value;
      return NH_OK;
    case NH_SET_SK_ROLES:
      NH_CHECK(value >= -1 && value <= 1, NH_ERR_ARG,
               "NH_SET_SK_ROLES: value %d not in -1..1", value);
      ctx->sk_roles = value;
      return NH_OK;
    case NH_SET_BLOCK_HV:
      NH_CHECK(value >= -1 && value <= 2, NH_ERR_ARG,
               "NH_SET_BLOCK_HV: value %d not in -1..2", value);
      ctx->block_hv = value;
      return NH_OK;
    case NH_SET_RING_STEPS:
      NH_CHECK(value == 0 || value == 8 || value == 16 || value == 32 || value == 64 ||
                   value == 128,
               NH_ERR_ARG, "NH_SET_RING_STEPS: value %d not 0, 8, 16, 32, 64 or 128",
               value);
      ctx->ring_steps = value;
      return NH_OK;
    default:
      NH_CHECK(false, NH_ERR_ARG, "nh_context_set: unknown setting %d", what);
  }
}

const char *nh_last_error(void) { return g_err; }

nh_status nh_context_create(const nh_model *model, int device,
                            nh_context **out) {
  return create(model, device, out);
}

void nh_context_destroy(nh_context *ctx) { destroy(ctx); }

int64_t nh_context_dim(const nh_context *ctx) { return ctx ? ctx->n : 0; }

nh_status nh_ham_build(nh_context *ctx, int layout, int precision) {
  NH_TRY(check_ctx(ctx));
  return assemble(ctx, layout, precision);
}

nh_status nh_ham_set_precision(nh_context *ctx, int precision) {
  NH_TRY(check_ctx(ctx));
  NH_HIP(hipSetDevice(ctx->device));
  return select_precision(ctx, precision);
}

nh_status nh_hv(nh_context *ctx, double shift, const nh_cplx *x, nh_cplx *y,
                int ptr_kind) {
  return nh_hV(ctx, shift, x, y, 1, ctx ? ctx->n : 0, ptr_kind);
}

nh_status nh_hV(nh_context *ctx, double shift, const nh_cplx *X, nh_cplx *Y,
                int ncol, int64_t ld, int ptr_kind) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(X && Y && X != Y, NH_ERR_ARG, "H·v: x and y must be distinct");
  NH_CHECK(ncol >= 1 && ld >= ctx->n, NH_ERR_ARG, "H·V: bad ncol/ld");
  if (ptr_kind == NH_PTR_DEVICE) {
    NH_TRY(launch_hv(ctx, shift, X, Y, ncol, ld));
    NH_HIP(hipStreamSynchronize(ctx->stream));
    return NH_OK;
  }
  NH_CHECK(ptr_kind == NH_PTR_HOST, NH_ERR_ARG, "bad ptr_kind");
  const int64_t n = ctx->n;
  NH_TRY(ensure_buffers(ctx, n * ncol));
  for (int k = 0; k < ncol; ++k)
    NH_HIP(hipMemcpyAsync(ctx->xbuf + k * n, X + k * ld, n * sizeof(nh_cplx),
                          hipMemcpyHostToDevice, ctx->stream));
  NH_TRY(launch_hv(ctx, shift, ctx->xbuf, ctx->ybuf, ncol, n));
  for (int k = 0; k < ncol; ++k)
    NH_HIP(hipMemcpyAsync(Y + k * ld, ctx->ybuf + k * n, n * sizeof(nh_cplx),
                          hipMemcpyDeviceToHost, ctx->stream));
  NH_HIP(hipStreamSynchronize(ctx->stream));
  return NH_OK;
}

nh_status nh_ham_export(nh_context *ctx, double *onsite, double *blocks) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(ctx->onsite && ctx->blocks64, NH_ERR_STATE,
           "nh_ham_export: needs a build in NH_LAYOUT_BLOCKS");
  const int A = ctx->n_atoms;
  if (onsite) {
    std::vector<double> tmp((size_t)A * kOnsiteStride);
    NH_HIP(hipMemcpy(tmp.data(), ctx->onsite, tmp.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < A; ++i)
      memcpy(onsite + (size_t)i * 17, tmp.data() + (size_t)i * kOnsiteStride,
             17 * sizeof(double));
  }
  if (blocks)
    NH_HIP(hipMemcpy(blocks, ctx->blocks64, (size_t)A * 400 * sizeof(double),
                     hipMemcpyDeviceToHost));
  return NH_OK;
}

nh_status nh_context_mem_info(nh_context *ctx, int64_t *free_bytes,
                              int64_t *pool_idle_bytes) {
  NH_TRY(check_ctx(ctx));
  size_t fr = 0, tot = 0;
  NH_HIP(hipMemGetInfo(&fr, &tot));
  if (free_bytes) *free_bytes = (int64_t)fr;
  if (pool_idle_bytes) *pool_idle_bytes = (int64_t)pool_free_bytes(ctx);
  return NH_OK;
}

nh_status nh_device_alloc(nh_context *ctx, int64_t bytes, void **dptr) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(dptr && bytes >= 0, NH_ERR_ARG, "nh_device_alloc: bad arguments");
  return device_malloc(ctx, dptr, (size_t)bytes);
}

nh_status nh_device_free(nh_context *ctx, void *dptr) {
  NH_TRY(check_ctx(ctx));
  NH_HIP(hipStreamSynchronize(ctx->stream));
  if (dptr) NH_HIP(hipFree(dptr));
  return NH_OK;
}

nh_status nh_memcpy_h2d(nh_context *ctx, void *dst, const void *src,
                        int64_t bytes) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(dst && src && bytes >= 0, NH_ERR_ARG, "nh_memcpy_h2d: bad arguments");
  NH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  NH_HIP(hipStreamSynchronize(ctx->stream));
  return NH_OK;
}

nh_status nh_memcpy_d2h(nh_context *ctx, void *dst, const void *src,
                        int64_t bytes) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(dst && src && bytes >= 0, NH_ERR_ARG, "nh_memcpy_d2h: bad arguments");
  NH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  NH_HIP(hipStreamSynchronize(ctx->stream));
  return NH_OK;
}

nh_status nh_dot(nh_context *ctx, int64_t n, const nh_cplx *x,
                 const nh_cplx *y, nh_cplx *out) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(x && y && out && n >= 0, NH_ERR_ARG, "nh_dot: bad arguments");
  double r[2];
  NH_TRY(reduce_dot(ctx, n, x, y, r));
  out->re = r[0];
  out->im = r[1];
  return NH_OK;
}

nh_status nh_dot_noalloc(nh_context *ctx, int64_t n, const nh_cplx *x,
                         const nh_cplx *y, nh_cplx *out) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(x && y && out && n >= 0, NH_ERR_ARG, "nh_dot_noalloc: bad arguments");
  double r[2];
  NH_TRY(reduce_dot_noalloc(ctx, n, x, y, r));
  out->re = r[0];
  out->im = r[1];
  return NH_OK;
}

nh_status nh_nrm2(nh_context *ctx, int64_t n, const nh_cplx *x, double *out) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(x && out && n >= 0, NH_ERR_ARG, "nh_nrm2: bad arguments");
  double s;
  NH_TRY(reduce_nrm2sq(ctx, n, x, &s));
  *out = sqrt(s);
  return NH_OK;
}

nh_status nh_axpby(nh_context *ctx, int64_t n, nh_cplx a, const nh_cplx *x,
                   nh_cplx b, nh_cplx *y) {
  NH_TRY(check_ctx(ctx));
  NH_CHECK(x && y && n >= 0, NH_ERR_ARG, "nh_axpby: bad arguments");
  NH_TRY(launch_axpby(ctx, n, a, x, b, y));
  NH_HIP(hipStreamSynchronize(ctx->stream));
  return NH_OK;
}

}  // extern "C"
