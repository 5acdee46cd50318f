// Internal state of libnemohip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nemohip.h"

namespace nh {

void set_error(const char *fmt, ...);

#define NH_HIP(call)                                                       \
  do {                                                                     \
    const hipError_t nh_hip_status_ = (call);                              \
    if (nh_hip_status_ != hipSuccess) {                                    \
      ::nh::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,          \
                      hipGetErrorString(nh_hip_status_));                  \
      return NH_ERR_HIP;                                                   \
    }                                                                      \
  } while (0)

#define NH_CHECK(cond, code, ...)                                          \
  do {                                                                     \
    if (!(cond)) {                                                         \
      ::nh::set_error(__VA_ARGS__);                                        \
      return code;                                                         \
    }                                                                      \
  } while (0)

// (the local's name must not collide with any identifier used in `expr`,
// e.g. a member s_, which it would shadow inside its own initialiser)
#define NH_TRY(expr)                                                       \
  do {                                                                     \
    const nh_status nh_try_status_ = (expr);                               \
    if (nh_try_status_ != NH_OK) return nh_try_status_;                    \
  } while (0)

// Onsite record per atom: 10 diagonal energies, 6 surface s-p couplings,
// spin-orbit lambda (17 doubles, padded to 18 for 16-byte alignment).
constexpr int kOnsiteStride = 18;
// Per-bond SK record: l, m, n and the 21 scaled integrals (24 values).
constexpr int kSkStride = 24;
// Matrix-free layout, per (cation species, anion species) pair: the 21
// unscaled two-centre integrals (parameters 10..30), the 15 strain
// exponents (strain parameters 0..14) and the unstrained lattice constant.
constexpr int kMfPair = 40;
constexpr int kMfEta = 21, kMfA0 = 36;

}  // namespace nh

struct nh_context {
  int device = 0;
  int nb = 0;        // 10 or 20
  int n_atoms = 0;
  int n_species = 0;
  int n_cations = 0;
  int n_anions = 0;
  int64_t n = 0;     // n_atoms * nb
  double lattice[3] = {0, 0, 0};
  int passivate = 0;
  double db_shift = 0;
  hipStream_t stream = nullptr;

  // model (device)
  double *parmat = nullptr;     // [S][S][35]
  double *strnmat = nullptr;    // [S][S][33]
  int32_t *species = nullptr;   // [A]
  int32_t *is_cation = nullptr; // [A]
  int32_t *nbr = nullptr;       // [A][4]
  int32_t *bond_q = nullptr;    // [A][4][3]
  double *disp = nullptr;       // [A][3]
  double *phi = nullptr;        // [A] or null
  int32_t *cations = nullptr;   // [Nc] atom index of each cation
  int32_t *anions = nullptr;    // [Na] atom index of each anion
  int32_t *bond_of = nullptr;   // [A][4] cation-owned bond id (ci*4+slot)
  int32_t *rev_slot = nullptr;  // [A][4] slot of this atom in the nbr's list

  // kernel variants (nh_context_set)
  // stored-block H·v: neighbour x staged in LDS (measured on the 786,520-
  // atom dot: faster for the plain product, slower inside the fused Lanczos
  // kernel, so off by default)
  int hv_staged = -1;  // NH_SET_HV_STAGED (-1: by layout, see use_staged)
  int ring_steps = 0;  // NH_SET_RING_STEPS (0: by free memory)
  int block_hv = -1;   // NH_SET_BLOCK_HV (-1: by measurement)
  int sk_roles = -1;   // NH_SET_SK_ROLES (-1: default, sk_mixed_roles)
  int sk_lanes = -1;   // NH_SET_SK_LANES (-1: default, sk_split_lanes)
  int lz_vectors = -1; // NH_SET_LZ_VECTORS (-1: default, lz_fp32_vectors)

  // assembled Hamiltonian
  int layout = -1;
  int precision = -1;
  double *onsite = nullptr;     // [A][kOnsiteStride]
  double *blocks64 = nullptr;   // [A][4][100] (NH_LAYOUT_BLOCKS, FP64)
  float *blocks32 = nullptr;    // [A][4][100] (NH_LAYOUT_BLOCKS, FP32)
  float *onsite32 = nullptr;    // [A][kOnsiteStride] (FP32 runs)
  double *sk64 = nullptr;       // [Nc*4][kSkStride] (NH_LAYOUT_SK)
  float *sk32 = nullptr;
  double *corr = nullptr;       // [A][4][10] Boykin corrections (scratch)
  // NH_LAYOUT_MATRIX_FREE: bond records recomputed on every product from
  double4 *mf_geo = nullptr;    // [A] displacement (nm) x, y, z; species
  int32_t *mf_q = nullptr;      // [A][4] ideal bond vector of the slot seen
                                // from its cation, units a/4, 3 x int8
  double *mf_pair = nullptr;    // [S][S][kMfPair] (cation, anion) constants

  // work buffers
  double *red = nullptr;        // reduction partials
  int red_cap = 0;
  double *hred = nullptr;       // pinned host mirror of reductions
  nh_cplx *xbuf = nullptr;      // host-pointer staging
  nh_cplx *ybuf = nullptr;
  int64_t buf_cap = 0;          // complex elements in xbuf/ybuf
  // solver workspace kept across solves: blocks a solver releases stay
  // allocated for the next request (a hipFree of tens of GB can stall for
  // a second or more); freed with the context
  struct PoolBlock {
    void *p;
    size_t bytes;
    bool used;
  };
  std::vector<PoolBlock> pool;
};

namespace nh {
// LDS staging of the neighbours' entries in H·v: measured faster for the SK
// layout (quad gather), slower for the stored-block layout inside the
// fused Lanczos step; NH_SET_HV_STAGED overrides.
// Layouts applied by the 4-lanes-per-atom bond kernels (nh_hv_sk.h): the
// stored SK records and the matrix-free one.
inline bool sk_like(const nh_context *c) {
  return c->layout == NH_LAYOUT_SK || c->layout == NH_LAYOUT_MATRIX_FREE;
}
inline bool use_staged(const nh_context *c) {
  return c->hv_staged < 0 ? sk_like(c) : c->hv_staged != 0;
}
// SK / matrix-free products: both roles in one workgroup (waves 0-1 the
// cations, waves 2-3 the anions at the same list position, so both rows of
// a bond read its record within one workgroup) or one role per workgroup.
// Default: both roles with FP64 products, one role with FP32 (measured on
// a 786,520-atom dot, fused Lanczos kernel: FP64 0.329 against 0.338 ms,
// FP32 0.230 against 0.229 ms)
inline bool sk_mixed_roles(const nh_context *c) {
  return c->sk_roles < 0 ? c->precision == NH_PREC_FP64 : c->sk_roles != 0;
}
// SK / matrix-free products of 20-band models with the LDS gather: 8 lanes
// per atom (one per bond and spin block) instead of 4 (one per bond).
// Default: 8 for matrix-free FP32 products only (measured on a 786,520-atom
// dot, fused Lanczos kernel: matrix-free FP32 0.301 against 0.345 ms; SK
// FP32 0.283 against 0.230, SK FP64 0.467 against 0.330, matrix-free FP64
// 0.487 against 0.334 ms)
inline bool sk_split_lanes(const nh_context *c) {
  if (c->nb != 20 || !use_staged(c)) return false;
  if (c->sk_lanes < 0)
    return c->layout == NH_LAYOUT_MATRIX_FREE && c->precision == NH_PREC_FP32;
  return c->sk_lanes == 8;
}
// Lanczos vectors stored in FP32 (FP32 SK or matrix-free products with the
// LDS gather only).  Default on (measured on a 786,520-atom dot, FP32 SK
// products: the same 7,901 steps to the same eigenvalues, fused H*v 0.208
// / 0.228 ms in the two passes against 0.229 / 0.301 ms, FP32 stage 4.73 s
// against 6.23 s, with a separate update kernel)
inline bool lz_fp32_vectors(const nh_context *c) {
  return sk_like(c) && c->precision == NH_PREC_FP32 && use_staged(c) &&
         c->lz_vectors != 0;
}
}  // namespace nh


namespace nh {
// Sampled kernel timing with HIP events on the launching stream: begin()/
// end() around a launch; collect() after a synchronisation folds the
// elapsed times into the mean.
struct KernelTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  double sum_ms = 0.0;
  int count = 0;
  hipEvent_t pending = nullptr;
  ~KernelTimer() { release(); }
  void release() {
    for (auto &e : ev) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    ev.clear();
    if (pending) (void)hipEventDestroy(pending);
    pending = nullptr;
  }
  nh_status begin(hipStream_t s) {
    NH_HIP(hipEventCreate(&pending));
    NH_HIP(hipEventRecord(pending, s));
    return NH_OK;
  }
  nh_status end(hipStream_t s) {
    hipEvent_t e;
    NH_HIP(hipEventCreate(&e));
    NH_HIP(hipEventRecord(e, s));
    ev.push_back({pending, e});
    pending = nullptr;
    return NH_OK;
  }
  nh_status collect(hipStream_t s) {
    NH_HIP(hipStreamSynchronize(s));
    for (auto &e : ev) {
      float ms = 0.f;
      NH_HIP(hipEventElapsedTime(&ms, e.first, e.second));
      sum_ms += ms;
      ++count;
    }
    release();
    return NH_OK;
  }
  double mean_ms() const { return count ? sum_ms / count : 0.0; }
};

// Progress lines on stderr for long solves (NEMOHIP_PROGRESS=1): at most
// one every 15 s, so that a caller's log shows the solve is alive.
struct Progress {
  bool on = false;
  double last = 0.0;
  Progress();
  void tick(const char *what, int iter, int extra);
};

// kernels / launchers (nh_assemble.hip, nh_hv.hip, nh_vector.hip)
nh_status assemble(nh_context *c, int layout, int precision);
nh_status select_precision(nh_context *c, int precision);
nh_status launch_hv(nh_context *c, double shift, const nh_cplx *x,
                    nh_cplx *y, int ncol, int64_t ld);
// block H·V of 8 vectors on the matrix cores (FP32 builds; nh_hv_mfma.hip)
bool hV_mfma_supported(const nh_context *c, int ncol);
nh_status launch_hV_mfma(nh_context *c, double shift, const nh_cplx *x,
                         nh_cplx *y, int64_t ld);
nh_status reduce_dot(nh_context *c, int64_t n, const nh_cplx *x,
                     const nh_cplx *y, double *out2 /* host re,im */);
nh_status reduce_nrm2sq(nh_context *c, int64_t n, const nh_cplx *x,
                        double *out);
// nh_dot_noalloc (nh_vector.hip): file-scope scratch per device, its address
// resolved once per device at context creation
nh_status init_noalloc_scratch(int device);
nh_status reduce_dot_noalloc(nh_context *c, int64_t n, const nh_cplx *x,
                             const nh_cplx *y, double *out2);
nh_status launch_axpby(nh_context *c, int64_t n, nh_cplx a, const nh_cplx *x,
                       nh_cplx b, nh_cplx *y);
nh_status ensure_reduce(nh_context *c, int blocks);
// hipMalloc that, when the device is full, frees the idle workspace blocks
// of every context on the device and tries once more (NH_ERR_NOMEM after
// that); every device allocation of the library goes through it
nh_status device_malloc(nh_context *c, void **p, size_t bytes);
template <typename T>
inline nh_status device_alloc(nh_context *c, T **p, size_t count) {
  return device_malloc(c, reinterpret_cast<void **>(p), count * sizeof(T));
}
// The reference's random starts on the host (nh_random.hip): glibc's
// rand() after srand(seed), LAPACK's dlaruv / zlarnv(idist = 1), and the
// Lanczos Random start (rand() / RAND_MAX, 0) built from them.
class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed);
  int32_t next() { return (int32_t)(step() >> 1); }

 private:
  uint32_t step();
  uint32_t r_[34];
  int k_ = 0;
};
void dlaruv(int32_t iseed[4], int n, double *x);
void zlarnv_uniform(int32_t iseed[4], int64_t n, nh_cplx *x);
void rand_start(uint32_t seed, int64_t n, nh_cplx *x);
// device memory from the context's workspace pool: the smallest free block
// of at least `bytes` and at most twice that, else a new allocation (after
// returning the free blocks to the device if it is full); pool_release
// hands a block back (work queued on the context's stream that uses it may
// still be running: the next user is ordered after it on that stream)
nh_status pool_alloc(nh_context *c, void **p, size_t bytes);
void pool_release(nh_context *c, void *p);
// bytes in the idle blocks of every context on c's device
size_t pool_free_bytes(const nh_context *c);
nh_status ensure_buffers(nh_context *c, int64_t elems);
}  // namespace nh
