// Device-resident Lanczos eigensolver (no reorthogonalisation), the
// counterpart of eigensolve_lanczos (NEMO_3D/eigen/src/lanczos.c:1823-2089).
//
// One iteration (single_lanczos_iteration, lanczos.c:421-757) is two fused
// kernels plus two one-workgroup reductions, all on one stream:
//   k_lz_hv     q = r / beta_{k-1} (written into a ring slot), y = H q, and
//               per-workgroup partials of alpha = Re<q, y>;
//   k_lz_update r = y - alpha q - beta_{k-1} q_prev with partials of |r|^2.
// With FP32 products and FP32 Lanczos vectors (SK and matrix-free layouts)
// it is one kernel and one reduction: the product forms r_k from u_{k-1}
// and r_{k-1} as it reads them, computes u_k = H q_k - beta_k q_{k-1} and
// the partials of alpha_k, |u_k|^2 and |q_k|^2 (k_lz_hv_sk, Paige's order).
// The reference's REDUCE_OPT variant of the iteration (lanczos.c:441-456)
// defers the vector update into the next iteration in the same spirit, to
// merge its two reductions into one.
// alpha and beta stay in device arrays; the host copies the tridiagonal
// every ConvCheckSkipRate iterations and analyses it (Sturm bisection in
// the Erange windows, convergence by the trailing components of the
// tridiagonal eigenvector, find_converged_eigval Method 2,
// lanczos.c:764-996).  Eigenvectors come from a second pass over the same
// recurrence (eigenvectorsLanczos, lanczos.c:1364-1689): the Lanczos
// vectors land in a ring (K = 128 steps, fewer when memory is short) and
// every K iterations one kernel adds Ring * Z to the eigenvector block.
// The second pass reuses alpha and beta, so with the SK layout its H*v
// kernel also applies the three-term step (no separate update kernel, no
// reductions), and the ring holds the r vectors themselves.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <thread>
#include <type_traits>
#include <vector>

#include "nh_hv_row.h"
#include "nh_hv_sk.h"
#include "nh_vector.h"

namespace nh {
namespace {

constexpr int kRingMax = 128;
constexpr int kMaxVec = 64;

__device__ __forceinline__ double wg_sum320(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __shared__ double s[kThreads / 64];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) s[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) t += s[k];
  }
  return t;
}

template <int NSP, typename TB, typename TA, bool STAGED>
__global__ __launch_bounds__(kThreads) void k_lz_hv(
    int A, const double *__restrict__ onsite, const TB *__restrict__ blocks,
    const int32_t *__restrict__ nbr, const double2 *__restrict__ R,
    double2 *__restrict__ Q, double2 *__restrict__ Y,
    const double *__restrict__ beta_prev, double *__restrict__ part) {
  constexpr int NB = NSP * 10;
  const int la = threadIdx.x / 10;
  const int atom = blockIdx.x * kAtomsPerBlock + la;
  const int o = threadIdx.x % 10;
  const double inv = beta_prev ? 1.0 / beta_prev[0] : 1.0;
  double local = 0.0;
  double acc_re[NSP], acc_im[NSP];
  double2 xi[NSP];
  if constexpr (STAGED) {
    __shared__ double2 sx[kAtomsPerBlock * 4 * NB];
    if (atom < A) stage_nbr_x<NSP>(atom, o, nbr, R, sx + la * 4 * NB);
    __syncthreads();
    if (atom < A)
      hv_row_staged<NSP, TB, TA>(atom, o, onsite, blocks, 0.0, R,
                                 sx + la * 4 * NB, acc_re, acc_im, xi);
  } else {
    if (atom < A)
      hv_row<NSP, TB, TA>(atom, o, onsite, blocks, nbr, 0.0, R, acc_re, acc_im,
                          xi);
  }
  if (atom < A) {
    const int64_t base = (int64_t)atom * NB;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
      const double2 q = make_double2(xi[s].x * inv, xi[s].y * inv);
      const double2 y = make_double2(acc_re[s] * inv, acc_im[s] * inv);
      if (Q) Q[base + s * 10 + o] = q;
      Y[base + s * 10 + o] = y;
      local += q.x * y.x + q.y * y.y;
    }
  }
  if (part) {
    const double t = wg_sum320(local);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
  }
}

// SK-layout variant of k_lz_hv (mapping of nh_hv_sk.h: 4 lanes per atom,
// cation and anion workgroups).
// Workgroup sum for the fused SK kernel: with the LDS gather the staging
// area is reused (the kernel's LDS stays at 80 KB for two groups per CU).
template <bool QUAD>
struct LzBlockSum {
  template <typename S>
  __device__ __forceinline__ static void run(double &a, double &b, S *) {
    block_sum2(a, b);
  }
};
template <>
struct LzBlockSum<true> {
  template <typename S>
  __device__ __forceinline__ static void run(double &a, double &b, S *L) {
    block_sum2_lds(a, b, reinterpret_cast<double *>(L->buf));
  }
};

// One element of the three-term step y - alpha q - beta_prev q_prev with
// q_prev = r_prev * (1 / beta_prev2); rp null on the first step.  The
// operations are spelled out (explicit fma, y kept opaque) so that the
// separate update kernel and the H*v kernel of the eigenvector pass, which
// fuses the step, round identically: the second pass must retrace the first
// pass's trajectory bit for bit.
__device__ __forceinline__ double2 lz_three_term_v(double2 y, double2 q,
                                                   bool has_p, double2 r,
                                                   double al, double be,
                                                   double invp) {
  asm volatile("" : "+v"(y.x), "+v"(y.y));
  if (has_p) {
    const double2 p = make_double2(r.x * invp, r.y * invp);
    y.x = y.x - __builtin_fma(al, q.x, be * p.x);
    y.y = y.y - __builtin_fma(al, q.y, be * p.y);
  } else {
    y.x = __builtin_fma(-al, q.x, y.x);
    y.y = __builtin_fma(-al, q.y, y.y);
  }
  return y;
}
// the same step with r_prev read from rp (null: first step; FP64 or FP32
// storage)
template <typename XO>
__device__ __forceinline__ double2 lz_three_term(double2 y, double2 q,
                                                 const XO *rp, double al,
                                                 double be, double invp) {
  return lz_three_term_v(y, q, rp != nullptr,
                         rp ? widen2(*rp) : make_double2(0.0, 0.0), al, be, invp);
}

// FP32 Lanczos vectors (Paige's order of the step): u_k = H q_k - beta_k
// q_{k-1} computed in FP64 and rounded to FP32, alpha_k = <q_k, u_k>,
// r_{k+1} = u_k - alpha_k q_k (lz_defer1).  y is kept opaque so that every
// kernel rounds u_k from the same two operations.
__device__ __forceinline__ float2 lz_u32(double2 y, double2 rp, double be, double invp) {
  asm volatile("" : "+v"(y.x), "+v"(y.y));
  return make_float2((float)__builtin_fma(-be, rp.x * invp, y.x),
                     (float)__builtin_fma(-be, rp.y * invp, y.y));
}
// per-workgroup partials of the FP32-vector eigenvalue pass: alpha_k,
// |u_k|^2, |q_k|^2 (beta_{k+1}^2 = |u|^2 - 2 alpha^2 + alpha^2 |q|^2)
struct LzPart3 {
  double a = 0.0, u = 0.0, q = 0.0;
  __device__ __forceinline__ void add(double2 qv, float2 uv) {
    const double ux = uv.x, uy = uv.y;
    a += qv.x * ux + qv.y * uy;
    u += ux * ux + uy * uy;
    q += qv.x * qv.x + qv.y * qv.y;
  }
  template <bool QUAD, typename S>
  __device__ __forceinline__ void store(double *part, S *L) {
    LzBlockSum<QUAD>::run(a, u, L);
    double z = 0.0;
    LzBlockSum<QUAD>::run(q, z, L);
    if (threadIdx.x == 0) {
      part[3 * blockIdx.x] = a;
      part[3 * blockIdx.x + 1] = u;
      part[3 * blockIdx.x + 2] = q;
    }
  }
};

// Epilogue staging of the FP32-vector kernels: the atom's own entries of
// r_k and r_{k-1} (FP32) and its onsite record
template <int NSP, int ATOMS = kSkAtomsPerBlock>
struct SkOwnStage32 {
  float2 x[ATOMS][NSP * 10];
  float2 rp[ATOMS][NSP * 10];
  double2 os[ATOMS][kOnsiteStride / 2];
};

// XO: storage of the Lanczos vectors r.  double2: FP64 r (own entries,
// r_prev) with an FP32 copy for the gathers of FP32 products.  float2
// (FP32 products, LDS gather): the vectors are FP32 and the step runs in
// Paige's order (lz_u32, lz_defer1).  Eigenvalue pass: Rv = u_{k-1}, Rp =
// r_{k-1}; r_k is formed on the fly (gather and own entries) and written
// to Rout, u_k to Y32, the partials of LzPart3 to part (no update
// kernel).  Eigenvector pass (FUSE): Rv = Rv32 = r_k, Rp = r_{k-1}, the
// rounded r_{k+1} to Y32.  The step's arithmetic stays FP64.
template <int NSP, typename R, bool QUAD, bool FUSE, typename SRC, typename XO = double2>
__global__ __launch_bounds__(kSkThreads, sizeof(R) == 4 ? 4 : 2) void k_lz_hv_sk(
    int nc, const int32_t *__restrict__ cations, int na,
    const int32_t *__restrict__ anions, SkGrid grid,
    const double *__restrict__ onsite, const SRC src,
    const int32_t *__restrict__ nbr, const XO *__restrict__ Rv, const float2 *__restrict__ Rv32,
    double2 *__restrict__ Q, double2 *__restrict__ Y,
    const double *__restrict__ beta_prev, double *__restrict__ part,
    const double *__restrict__ alpha, const XO *__restrict__ Rp,
    const double *__restrict__ beta_prev2, float2 *__restrict__ Y32,
    float2 *__restrict__ Rout, const double *__restrict__ alpha_prev) {
  constexpr int NB = NSP * 10;
  constexpr bool V32 = std::is_same<XO, float2>::value;
  constexpr bool DEFER = V32 && !FUSE;
  static_assert(!V32 || (sizeof(R) == 4 && QUAD),
                "FP32 Lanczos vectors need FP32 products and the LDS gather");
  constexpr int NO = (NB + 3) / 4;
  const SkLane ln = sk_lane_of(grid, blockIdx.x, threadIdx.x);
  // the role is uniform across the wavefront; read as a scalar, the role
  // branches below are scalar branches (one per wave)
  const bool cat = __builtin_amdgcn_readfirstlane((int)ln.cat) != 0;
  const int li = ln.li;
  const int slot = threadIdx.x & 3;
  const bool active = li < (cat ? nc : na);
  const int atom = active ? (cat ? cations[li] : anions[li]) : 0;
  const double inv = beta_prev ? 1.0 / beta_prev[0] : 1.0;
  // FUSE (eigenvector pass): alpha is known, so the three-term step of
  // k_lz_update is applied here and Y receives r_{k+1}
  const double al = FUSE ? alpha[0] : 0.0;
  const double be = beta_prev ? beta_prev[0] : 0.0;
  const double invp = beta_prev2 ? 1.0 / beta_prev2[0] : 1.0;
  double own_re[NO], own_im[NO];
  // FP32 products gather from the FP32 copy of r (the same rounding as
  // converting the FP64 entries on load)
  using XT = typename std::conditional<sizeof(R) == 4, float2, double2>::type;
  const XT *xg;
  if constexpr (sizeof(R) == 4) xg = Rv32;
  else xg = Rv;
  // DEFER: r_k = lz_defer1(u_{k-1}, r_{k-1}, alpha_{k-1} / beta_{k-1});
  // FUSE: r_{k+1} = lz_defer1(u_k, r_k, alpha_k / beta_k)
  const double alp = DEFER && alpha_prev ? alpha_prev[0] : 0.0;
  const LzCoef ap = lz_defer_coef(alp, invp), af = lz_defer_coef(al, inv);
  LzDeferSrc dsrc{nullptr, ap};
  if constexpr (DEFER) dsrc.rp = Rp;
  SkQuadStage<NSP, XT> *L = nullptr;
  if constexpr (QUAD) {
    __shared__ SkQuadStage<NSP, XT> stage;
    L = &stage;
  }
  if (cat)
    sk_bonds<NSP, R, true, XT, QUAD, DEFER>(atom, slot, active, src, nbr, xg, own_re,
                                            own_im, L, dsrc);
  else
    sk_bonds<NSP, R, false, XT, QUAD, DEFER>(atom, slot, active, src, nbr, xg, own_re,
                                             own_im, L, dsrc);
  double local = 0.0;
  LzPart3 p3;
  if constexpr (V32) {
    // FP32 vectors: the atom's own entries of r_k (DEFER: of u_{k-1}) and
    // r_{k-1} are requested in 16-byte pieces of two entries (lane `slot`
    // takes pieces slot, slot + 4, ...), with the onsite record, before the
    // epilogue's barriers, and staged in FP32 in the freed LDS; each lane
    // then finishes its rows 4 i + slot from there
    static_assert(sizeof(SkOwnStage32<NSP>) <= sizeof(SkQuadStage<NSP, XT>),
                  "epilogue staging must fit the gather staging");
    auto &S = *reinterpret_cast<SkOwnStage32<NSP> *>(L->buf);
    const int la = threadIdx.x / 4;
    constexpr int NC = NB / 2, NCL = (NC + 3) / 4;
    constexpr int KOS = SkOwnRegs<NSP>::kOs;
    float4 cx[NCL], cp[NCL];
    double2 co[KOS];
#pragma unroll
    for (int i = 0; i < NCL; ++i) cx[i] = cp[i] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int i = 0; i < KOS; ++i) co[i] = make_double2(0.0, 0.0);
    if (active) {
      const float4 *xs = reinterpret_cast<const float4 *>(Rv + (int64_t)atom * NB);
      const float4 *ps = reinterpret_cast<const float4 *>(Rp + (int64_t)atom * NB);
#pragma unroll
      for (int i = 0; i < NCL; ++i) {
        const int c = 4 * i + slot;
        if (c < NC) {
          cx[i] = xs[c];
          cp[i] = ps[c];
        }
      }
      const double2 *o2 = reinterpret_cast<const double2 *>(onsite + (size_t)atom * kOnsiteStride);
#pragma unroll
      for (int i = 0; i < KOS; ++i) {
        const int c = 4 * i + slot;
        if (c < kOnsiteStride / 2) co[i] = o2[c];
      }
      if constexpr (DEFER) {
        // r_k from u_{k-1} and r_{k-1}, written out for the next step
#pragma unroll
        for (int i = 0; i < NCL; ++i) {
          const int c = 4 * i + slot;
          if (c < NC) {
            cx[i] = make_float4(lz_defer1(cx[i].x, cp[i].x, ap), lz_defer1(cx[i].y, cp[i].y, ap),
                                lz_defer1(cx[i].z, cp[i].z, ap), lz_defer1(cx[i].w, cp[i].w, ap));
            if (Rout) reinterpret_cast<float4 *>(Rout + (int64_t)atom * NB)[c] = cx[i];
          }
        }
      }
    }
    __syncthreads();  // every lane has read its staged neighbour entries
    if (active) {
#pragma unroll
      for (int i = 0; i < NCL; ++i) {
        const int c = 4 * i + slot;
        if (c < NC) {
          S.x[la][2 * c] = make_float2(cx[i].x, cx[i].y);
          S.x[la][2 * c + 1] = make_float2(cx[i].z, cx[i].w);
          S.rp[la][2 * c] = make_float2(cp[i].x, cp[i].y);
          S.rp[la][2 * c + 1] = make_float2(cp[i].z, cp[i].w);
        }
      }
#pragma unroll
      for (int i = 0; i < KOS; ++i) {
        const int c = 4 * i + slot;
        if (c < kOnsiteStride / 2) S.os[la][c] = co[i];
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int i = 0; i < NO; ++i) {
        const int r = 4 * i + slot;
        if (r >= NB) break;
        double re = own_re[i], im = own_im[i];
        onsite_row<NSP>(reinterpret_cast<const double *>(S.os[la]), S.x[la],
                        r / 10, r % 10, 0.0, re, im);
        const float2 xf = S.x[la][r];
        const double2 q = make_double2((double)xf.x * inv, (double)xf.y * inv);
        const int64_t e = (int64_t)atom * NB + r;
        const double2 y = make_double2(re * inv, im * inv);
        const float2 u = lz_u32(y, widen2(S.rp[la][r]), be, invp);
        if constexpr (FUSE) {
          Y32[e] = make_float2(lz_defer1(u.x, xf.x, af), lz_defer1(u.y, xf.y, af));
        } else {
          Y32[e] = u;
          p3.add(q, u);
        }
      }
    }
  } else if constexpr (QUAD) {
    static_assert(sizeof(SkOwnStage<NSP>) <= sizeof(SkQuadStage<NSP, XT>),
                  "epilogue staging must fit the gather staging");
    auto &S = *reinterpret_cast<SkOwnStage<NSP> *>(L->buf);
    const int la = threadIdx.x / 4;
    // own entries, onsite record and (FUSE) r_prev of the lane's rows are
    // loaded before the epilogue's barriers, so the loads overlap them; the
    // first two are staged in the freed LDS after the first barrier
    SkOwnRegs<NSP> own;
    if (active)
      sk_load_own<NSP>(slot, Rv + (int64_t)atom * NB,
                       onsite + (size_t)atom * kOnsiteStride, own);
    double2 rpv[NO];
#pragma unroll
    for (int i = 0; i < NO; ++i) rpv[i] = make_double2(0.0, 0.0);
    if constexpr (FUSE) {
      if (active && Rp) {
#pragma unroll
        for (int i = 0; i < NO; ++i)
          if (4 * i + slot < NB) rpv[i] = widen2(Rp[(int64_t)atom * NB + 4 * i + slot]);
      }
    }
    __syncthreads();  // every lane has read its staged neighbour entries
    if (active) sk_store_own<NSP>(la, slot, own, S);
    __syncthreads();
    if (active) {
#pragma unroll
      for (int i = 0; i < NO; ++i) {
        const int r = 4 * i + slot;
        if (r >= NB) break;
        double re = own_re[i], im = own_im[i];
        onsite_row<NSP>(reinterpret_cast<const double *>(S.os[la]), S.x[la],
                        r / 10, r % 10, 0.0, re, im);
        const double2 q = make_double2(own.x[i].x * inv, own.x[i].y * inv);
        const int64_t e = (int64_t)atom * NB + r;
        double2 y = make_double2(re * inv, im * inv);
        if constexpr (FUSE) {
          y = lz_three_term_v(y, q, Rp != nullptr, rpv[i], al, be, invp);
          if (Y32) Y32[e] = make_float2((float)y.x, (float)y.y);
        }
        if (Q) Q[e] = q;
        Y[e] = y;
        local += q.x * y.x + q.y * y.y;
      }
    }
  } else if (active) {
    const double *os = onsite + (size_t)atom * kOnsiteStride;
    const XO *xa = Rv + (int64_t)atom * NB;
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      const int r = 4 * i + slot;
      if (r >= NB) break;
      double re = own_re[i], im = own_im[i];
      onsite_row<NSP>(os, xa, r / 10, r % 10, 0.0, re, im);
      const double2 xr = widen2(xa[r]);
      const double2 q = make_double2(xr.x * inv, xr.y * inv);
      const int64_t e = (int64_t)atom * NB + r;
      double2 y = make_double2(re * inv, im * inv);
      if constexpr (FUSE) {
        y = lz_three_term(y, q, Rp ? Rp + e : nullptr, al, be, invp);
        if (Y32) Y32[e] = make_float2((float)y.x, (float)y.y);
      }
      if (Q) Q[e] = q;
      Y[e] = y;
      local += q.x * y.x + q.y * y.y;
    }
  }
  if (part) {
    if constexpr (DEFER) {
      p3.store<QUAD>(part, L);
    } else {
      double z = 0.0;
      LzBlockSum<QUAD>::run(local, z, L);  // QUAD: reuses the staging LDS
      if (threadIdx.x == 0) part[blockIdx.x] = local;
    }
  }
}

// k_lz_hv_sk with 8 lanes per atom of a 20-band model (SkSplit: lane =
// spin block and bond), LDS gather; the same epilogue with lane (spin s,
// slot) finishing rows s * 10 + 4 i + slot.  Fewer vector registers per
// lane than k_lz_hv_sk (one spin block of the neighbour and of the
// accumulators), so more waves fit on a SIMD.
template <typename R, bool FUSE, typename SRC, typename XO = double2>
__global__ __launch_bounds__(kSkThreads, sizeof(R) == 4 ? 5 : 3) void k_lz_hv_sk_split(
    int nc, const int32_t *__restrict__ cations, int na,
    const int32_t *__restrict__ anions, SkGrid grid,
    const double *__restrict__ onsite, const SRC src,
    const int32_t *__restrict__ nbr, const XO *__restrict__ Rv, const float2 *__restrict__ Rv32,
    double2 *__restrict__ Q, double2 *__restrict__ Y,
    const double *__restrict__ beta_prev, double *__restrict__ part,
    const double *__restrict__ alpha, const XO *__restrict__ Rp,
    const double *__restrict__ beta_prev2, float2 *__restrict__ Y32,
    float2 *__restrict__ Rout, const double *__restrict__ alpha_prev) {
  constexpr int NSP = 2, NB = 20;
  constexpr bool V32 = std::is_same<XO, float2>::value;
  constexpr bool DEFER = V32 && !FUSE;
  static_assert(!V32 || sizeof(R) == 4, "FP32 Lanczos vectors need FP32 products");
  constexpr int kLanes = SkSplit::kLanes, kAtoms = SkSplit::kAtoms;
  const SkLane ln = sk_lane_of<kLanes>(grid, blockIdx.x, threadIdx.x);
  // the role is uniform across the wavefront (scalar role branches)
  const bool cat = __builtin_amdgcn_readfirstlane((int)ln.cat) != 0;
  const int li = ln.li;
  const int lane = threadIdx.x & (kLanes - 1), slot = lane & 3, spin = lane >> 2;
  const bool active = li < (cat ? nc : na);
  const int atom = active ? (cat ? cations[li] : anions[li]) : 0;
  const double inv = beta_prev ? 1.0 / beta_prev[0] : 1.0;
  const double al = FUSE ? alpha[0] : 0.0;
  const double be = beta_prev ? beta_prev[0] : 0.0;
  const double invp = beta_prev2 ? 1.0 / beta_prev2[0] : 1.0;
  double own_re[SkSplit::kOwn], own_im[SkSplit::kOwn];
  using XT = typename std::conditional<sizeof(R) == 4, float2, double2>::type;
  const XT *xg;
  if constexpr (sizeof(R) == 4) xg = Rv32;
  else xg = Rv;
  const double alp = DEFER && alpha_prev ? alpha_prev[0] : 0.0;
  const LzCoef ap = lz_defer_coef(alp, invp), af = lz_defer_coef(al, inv);
  LzDeferSrc dsrc{nullptr, ap};
  if constexpr (DEFER) dsrc.rp = Rp;
  __shared__ SkQuadStage<1, XT> stage;
  if (cat)
    sk_bonds_split<R, true, XT, DEFER>(atom, slot, spin, active, src, nbr, xg, own_re,
                                       own_im, stage, dsrc);
  else
    sk_bonds_split<R, false, XT, DEFER>(atom, slot, spin, active, src, nbr, xg, own_re,
                                        own_im, stage, dsrc);
  static_assert(sizeof(SkOwnStage<NSP, kAtoms>) <= sizeof(SkQuadStage<1, XT>),
                "epilogue staging must fit the gather staging");
  auto &S = *reinterpret_cast<SkOwnStage<NSP, kAtoms> *>(stage.buf);
  const int la = threadIdx.x / kLanes;
  SkOwnRegs<NSP, kLanes> own;
  if (active)
    sk_load_own<NSP, kLanes>(lane, Rv + (int64_t)atom * NB,
                             onsite + (size_t)atom * kOnsiteStride, own);
  double2 rpv[SkSplit::kOwn];
#pragma unroll
  for (int i = 0; i < SkSplit::kOwn; ++i) rpv[i] = make_double2(0.0, 0.0);
  if constexpr (FUSE || DEFER) {
    if (active && Rp) {
#pragma unroll
      for (int i = 0; i < SkSplit::kOwn; ++i)
        if (4 * i + slot < 10)
          rpv[i] = widen2(Rp[(int64_t)atom * NB + spin * 10 + 4 * i + slot]);
    }
  }
  if constexpr (DEFER) {
    // own entries of r_k (entries 8 i + lane, as loaded) from u_{k-1} and
    // r_{k-1}; written out for the next step
    if (active) {
#pragma unroll
      for (int i = 0; i < (NB + kLanes - 1) / kLanes; ++i) {
        const int r = kLanes * i + lane;
        if (r >= NB) break;
        const float2 rq = Rp[(int64_t)atom * NB + r];
        const float2 rk = make_float2(lz_defer1((float)own.x[i].x, rq.x, ap),
                                      lz_defer1((float)own.x[i].y, rq.y, ap));
        own.x[i] = widen2(rk);
        if (Rout) Rout[(int64_t)atom * NB + r] = rk;
      }
    }
  }
  __syncthreads();  // every lane has read its staged neighbour entries
  if (active) sk_store_own<NSP, kLanes, kAtoms>(la, lane, own, S);
  __syncthreads();
  double local = 0.0;
  LzPart3 p3;
  if (active) {
#pragma unroll
    for (int i = 0; i < SkSplit::kOwn; ++i) {
      const int c = 4 * i + slot;
      if (c >= 10) break;
      const int r = spin * 10 + c;
      double re = own_re[i], im = own_im[i];
      onsite_row<NSP>(reinterpret_cast<const double *>(S.os[la]), S.x[la], spin, c,
                      0.0, re, im);
      const double2 xo = S.x[la][r];
      const double2 q = make_double2(xo.x * inv, xo.y * inv);
      const int64_t e = (int64_t)atom * NB + r;
      double2 y = make_double2(re * inv, im * inv);
      if constexpr (V32) {
        const float2 u = lz_u32(y, rpv[i], be, invp);
        if constexpr (FUSE) {
          Y32[e] = make_float2(lz_defer1(u.x, (float)xo.x, af),
                               lz_defer1(u.y, (float)xo.y, af));
        } else {
          Y32[e] = u;
          p3.add(q, u);
        }
      } else {
        if constexpr (FUSE) {
          y = lz_three_term_v(y, q, Rp != nullptr, rpv[i], al, be, invp);
          if (Y32) Y32[e] = make_float2((float)y.x, (float)y.y);
        }
        if (Q) Q[e] = q;
        Y[e] = y;
        local += q.x * y.x + q.y * y.y;
      }
    }
  }
  if (part) {
    if constexpr (DEFER) {
      p3.store<true>(part, &stage);
    } else {
      double z = 0.0;
      LzBlockSum<true>::run(local, z, &stage);
      if (threadIdx.x == 0) part[blockIdx.x] = local;
    }
  }
}

// y <- y - alpha q - beta_prev q_prev with q = r * (1 / beta_prev) and
// q_prev = r_prev * (1 / beta_prev2) formed on the fly (the same products
// as the fused H·v kernel's q, so the trajectory does not depend on whether
// q was stored), partial sums of |y|^2.  Rp may be null (first step), and
// beta_prev / beta_prev2 null mean 1 (the start vector is normalised).
// Y32 (optional): FP32 copy of the new r for the next FP32 product.
__global__ __launch_bounds__(kRedThreads) void k_lz_update(
    int64_t n, double2 *__restrict__ Y, const double2 *__restrict__ Rk,
    const double2 *__restrict__ Rp, const double *__restrict__ alpha,
    const double *__restrict__ beta_prev, const double *__restrict__ beta_prev2,
    double *__restrict__ part, float2 *__restrict__ Y32) {
  const double al = alpha[0];
  const double be = beta_prev ? beta_prev[0] : 0.0;
  const double inv = beta_prev ? 1.0 / beta_prev[0] : 1.0;
  const double invp = beta_prev2 ? 1.0 / beta_prev2[0] : 1.0;
  double s = 0.0, z = 0.0;
  for (int64_t k = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * kRedThreads) {
    const double2 r = Rk[k];
    const double2 q = make_double2(r.x * inv, r.y * inv);
    const double2 y = lz_three_term(Y[k], q, Rp ? Rp + k : nullptr, al, be, invp);
    Y[k] = y;
    if (Y32) Y32[k] = make_float2((float)y.x, (float)y.y);
    s += y.x * y.x + y.y * y.y;
  }
  if (part) {
    block_sum2(s, z);
    if (threadIdx.x == 0) part[2 * blockIdx.x] = s;
  }
}

// out[0] = sum of nparts partials (stride `stride`); sqrt if take_sqrt.
__global__ __launch_bounds__(kRedThreads) void k_lz_scalar(
    int nparts, int stride, const double *__restrict__ part,
    double *__restrict__ out, int take_sqrt) {
  // eight independent partial sums per thread (all loads of a trip in
  // flight together), combined in a fixed order
  double p[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int k = threadIdx.x;
  for (; k + 7 * kRedThreads < nparts; k += 8 * kRedThreads) {
#pragma unroll
    for (int u = 0; u < 8; ++u) p[u] += part[(k + u * kRedThreads) * stride];
  }
  for (int u = 0; k < nparts; k += kRedThreads, ++u) p[u & 7] += part[k * stride];
  double a = ((p[0] + p[1]) + (p[2] + p[3])) + ((p[4] + p[5]) + (p[6] + p[7])), b = 0.0;
  block_sum2(a, b);
  if (threadIdx.x == 0) out[0] = take_sqrt ? sqrt(a) : a;
}

// FP32 Lanczos vectors: alpha_k and beta_{k+1} from the LzPart3 partials
// of nparts workgroups (fixed summation order):
// beta_{k+1}^2 = |u_k - alpha_k q_k|^2 = |u|^2 - 2 alpha^2 + alpha^2 |q|^2
__global__ __launch_bounds__(kRedThreads) void k_lz_scalars3(
    int nparts, const double *__restrict__ part, double *__restrict__ alpha,
    double *__restrict__ beta) {
  // four independent partial sums per component and thread (the loads of
  // a trip in flight together; constant indices keep them in registers),
  // combined in a fixed order
  double p[4][3] = {};
  int k = threadIdx.x;
  for (; k + 3 * kRedThreads < nparts; k += 4 * kRedThreads) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) p[u][c] += part[3 * (k + u * kRedThreads) + c];
  }
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (k + u * kRedThreads < nparts)
#pragma unroll
      for (int c = 0; c < 3; ++c) p[u][c] += part[3 * (k + u * kRedThreads) + c];
  double s[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) s[c] = (p[0][c] + p[1][c]) + (p[2][c] + p[3][c]);
  __shared__ double w[3][kRedThreads / 64];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s[c] += __shfl_down(s[c], off, 64);
    if ((threadIdx.x & 63) == 0) w[c][threadIdx.x >> 6] = s[c];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < kRedThreads / 64; ++k) t[c] += w[c][k];
    const double a = t[0];
    alpha[0] = a;
    beta[0] = sqrt(fmax(t[1] - 2.0 * a * a + a * a * t[2], 0.0));
  }
}

// V[e] += sum_k Z[e][k] Ring[k] for e < ne, k < K <= kRingMax (ring of
// FP64 q or of the FP32 copies of r).  Eight eigenvectors at a time stay in
// registers while the ring streams past in groups of eight slots.
template <typename T>
__global__ __launch_bounds__(256) void k_lz_accum(
    int64_t n, int ne, int K, const double *__restrict__ Z,
    const T *__restrict__ ring, int64_t ring_ld, double2 *__restrict__ V,
    int64_t v_ld) {
  __shared__ double zs[8 * kRingMax];
  const int K8 = (K + 7) / 8 * 8;
  for (int e0 = 0; e0 < ne; e0 += 8) {
    const int nb = min(8, ne - e0);
    __syncthreads();
    for (int t = threadIdx.x; t < 8 * K8; t += blockDim.x) {
      const int e = t / K8, k = t % K8;
      zs[t] = e < nb && k < K ? Z[(e0 + e) * K + k] : 0.0;
    }
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
      double2 v[8];
#pragma unroll
      for (int e = 0; e < 8; ++e)
        v[e] = e < nb ? V[(int64_t)(e0 + e) * v_ld + i] : make_double2(0.0, 0.0);
      for (int k0 = 0; k0 < K8; k0 += 8) {
        double2 q[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          q[j] = make_double2(0.0, 0.0);
          if (k0 + j < K) {
            const T r = ring[(int64_t)(k0 + j) * ring_ld + i];
            q[j] = make_double2(r.x, r.y);
          }
        }
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const double z = zs[e * K8 + k0 + j];
            v[e].x += z * q[j].x;
            v[e].y += z * q[j].y;
          }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (e < nb) V[(int64_t)(e0 + e) * v_ld + i] = v[e];
    }
  }
}

// The FP32 ring (copies of r) accumulated two rows per thread, so each
// ring load is 16 bytes; n even.
__global__ __launch_bounds__(256) void k_lz_accum32(
    int64_t n, int ne, int K, const double *__restrict__ Z,
    const float2 *__restrict__ ring, int64_t ring_ld, double2 *__restrict__ V,
    int64_t v_ld) {
  __shared__ double zs[8 * kRingMax];
  const int K8 = (K + 7) / 8 * 8;
  const int64_t n2 = n / 2;
  const float4 *r4 = reinterpret_cast<const float4 *>(ring);
  const int64_t ld4 = ring_ld / 2;
  for (int e0 = 0; e0 < ne; e0 += 8) {
    const int nb = min(8, ne - e0);
    __syncthreads();
    for (int t = threadIdx.x; t < 8 * K8; t += blockDim.x) {
      const int e = t / K8, k = t % K8;
      zs[t] = e < nb && k < K ? Z[(e0 + e) * K + k] : 0.0;
    }
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2;
         i += (int64_t)gridDim.x * blockDim.x) {
      double2 v0[8], v1[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        v0[e] = v1[e] = make_double2(0.0, 0.0);
        if (e < nb) {
          v0[e] = V[(int64_t)(e0 + e) * v_ld + 2 * i];
          v1[e] = V[(int64_t)(e0 + e) * v_ld + 2 * i + 1];
        }
      }
      for (int k0 = 0; k0 < K8; k0 += 8) {
        float4 q[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
          q[j] = k0 + j < K ? r4[(int64_t)(k0 + j) * ld4 + i] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int e = 0; e < 8; ++e)
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const double z = zs[e * K8 + k0 + j];
            v0[e].x += z * (double)q[j].x;
            v0[e].y += z * (double)q[j].y;
            v1[e].x += z * (double)q[j].z;
            v1[e].y += z * (double)q[j].w;
          }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (e < nb) {
          V[(int64_t)(e0 + e) * v_ld + 2 * i] = v0[e];
          V[(int64_t)(e0 + e) * v_ld + 2 * i + 1] = v1[e];
        }
    }
  }
}

__global__ void k_scale(int64_t n, double2 *__restrict__ x, double s) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    x[k].x *= s;
    x[k].y *= s;
  }
}

// ---------------- host-side tridiagonal analysis ---------------------------

int sturm(const std::vector<double> &d, const std::vector<double> &e, int m,
          double x) {
  int cnt = 0;
  double q = d[0] - x;
  if (q < 0) ++cnt;
  for (int i = 1; i < m; ++i) {
    if (q == 0.0) q = 1e-300;
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

// k-th smallest eigenvalue (0-based) of T_m by bisection on [lo, hi].
double bisect(const std::vector<double> &d, const std::vector<double> &e,
              int m, int k, double lo, double hi, double abstol) {
  for (int it = 0; it < 200 && hi - lo > abstol; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm(d, e, m, mid) > k) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Eigenvector of T_m at lambda by inverse iteration (tridiagonal Gaussian
// elimination with partial pivoting), from a fixed pseudo-random start.
void tri_eigvec(const std::vector<double> &d, const std::vector<double> &e,
                int m, double lambda, std::vector<double> &x) {
  x.assign(m, 0.0);
  uint64_t s = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < m; ++i) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    x[i] = 0.5 + (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
  const double pert = 1e-13 * (std::fabs(lambda) + 1.0);
  std::vector<double> a(m), b(m), c(m), c2(m), rhs(m);
  std::vector<char> swapped(m);
  for (int sweep = 0; sweep < 3; ++sweep) {
    // (T - lambda') x_new = x with a tridiagonal LU (a: sub, b: diag, c: sup)
    for (int i = 0; i < m; ++i) {
      b[i] = d[i] - lambda - pert;
      a[i] = i > 0 ? e[i - 1] : 0.0;
      c[i] = i + 1 < m ? e[i] : 0.0;
      c2[i] = 0.0;
      rhs[i] = x[i];
    }
    for (int i = 0; i + 1 < m; ++i) {
      // rows i and i+1; eliminate a[i+1]
      if (std::fabs(a[i + 1]) > std::fabs(b[i])) {
        std::swap(b[i], a[i + 1]);
        std::swap(c[i], b[i + 1]);
        std::swap(c2[i], c[i + 1]);
        std::swap(rhs[i], rhs[i + 1]);
      }
      if (b[i] == 0.0) b[i] = 1e-300;
      const double f = a[i + 1] / b[i];
      b[i + 1] -= f * c[i];
      c[i + 1] -= f * c2[i];
      rhs[i + 1] -= f * rhs[i];
    }
    if (b[m - 1] == 0.0) b[m - 1] = 1e-300;
    x[m - 1] = rhs[m - 1] / b[m - 1];
    if (m > 1) x[m - 2] = (rhs[m - 2] - c[m - 2] * x[m - 1]) / b[m - 2];
    for (int i = m - 3; i >= 0; --i)
      x[i] = (rhs[i] - c[i] * x[i + 1] - c2[i] * x[i + 2]) / b[i];
    double nrm = 0.0;
    for (int i = 0; i < m; ++i) nrm += x[i] * x[i];
    nrm = std::sqrt(nrm);
    for (int i = 0; i < m; ++i) x[i] /= nrm;
  }
}

// f(j) for j < count on up to 8 host threads (independent j)
template <typename F>
void host_parallel_for(int count, F f) {
  const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
  const int nt = std::min({count, hw, 8});
  if (nt <= 1) {
    for (int j = 0; j < count; ++j) f(j);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&f, t, nt, count] {
      for (int j = t; j < count; j += nt) f(j);
    });
  for (auto &th : pool) th.join();
}

struct Converged {
  double value;
  int iter;
  int window;
};

class LanczosRun {
 public:
  LanczosRun(nh_context *c) : c_(c) {}
  ~LanczosRun() {
    release_events();
    for (void *p : allocs_) pool_release(c_, p);
    if (pin_) (void)hipHostFree(pin_);
    if (fetch_ev_) (void)hipEventDestroy(fetch_ev_);
  }
  nh_status run(const nh_lanczos_opts *o, nh_eig_result *res);

 private:
  // from the context's workspace pool, handed back when the run ends
  nh_status alloc(void **p, size_t bytes) {
    NH_TRY(pool_alloc(c_, p, bytes));
    allocs_.push_back(*p);
    return NH_OK;
  }
  nh_status iterate(int k, bool with_reductions);
  nh_status collect_timing();
  nh_status fetch(int m);
  nh_status analyse(int m, const nh_lanczos_opts *o);
  bool window_done(int w, int req, int rule) const;

  nh_context *c_;
  std::vector<void *> allocs_;
  int64_t n_ = 0;
  double2 *R_ = nullptr, *Y_ = nullptr, *Rm1_ = nullptr, *ring_ = nullptr;
  // FP32 copies of r for the FP32 SK products (null otherwise)
  float2 *R32_ = nullptr, *Y32_ = nullptr;
  // FP32 Lanczos vectors (lz_fp32_vectors): the eigenvalue pass keeps
  // r_{k-1} in R32_ and u_{k-1} in U32_ and writes r_k, u_k to Y32_, UN32_
  // (step k = one launch plus k_lz_scalars3); R_ and Rm1_ are not allocated
  bool v32_ = false;
  float2 *U32_ = nullptr, *UN32_ = nullptr;
  // ring length of the eigenvector pass (<= kRingMax)
  int ring_k_ = kRingMax;
  // eigenvector pass with FP32 products: 2 K FP32 copies of r, slot
  // k % (2 K) holding r_k; the gather of step k reads slot k and writes
  // r_{k+1} to slot k + 1, and the accumulation reads them as its ring
  // (1 / beta_{k-1} folded into the coefficients)
  float2 *ring32_ = nullptr;
  // eigenvector pass with FP64 SK products: the r vectors themselves live in
  // a ring of 2 K FP64 slots (slot k % (2 K) holds r_k: read as own entries
  // and gathered at step k, r_(k-1) from slot k - 1, r_(k+1) written to slot
  // k + 1), so no q is stored; the accumulation reads them like the FP32 ring
  double2 *ringr_ = nullptr;
  nh_status build_start(const nh_lanczos_opts *o);
  nh_status upload_start(float2 *f32);
  double2 *R0_ = nullptr;  // the normalised start vector (device)
  double *alpha_ = nullptr, *beta_ = nullptr, *part_ = nullptr;
  int nparts_ = 0;
  std::vector<double> ha_, hb_;
  // pinned staging of alpha (first half) and beta, filled up to fetched_
  // by copies behind fetch_ev_
  double *pin_ = nullptr;
  int pin_cap_ = 0, fetched_ = 0;
  hipEvent_t fetch_ev_ = nullptr;
  std::vector<Converged> conv_;
  std::vector<double> lastRitz_[2];
  std::vector<char> lastConv_[2];
  // sampled H·v timing: one event pair every kTimeEvery iterations, per
  // pass (eigenvalue pass, eigenvector pass)
  static constexpr int kTimeEvery = 16;
  struct Timed {
    hipEvent_t start, stop;
    bool value_pass;
  };
  std::vector<Timed> ev_;
  double hv_ms_sum_[2] = {0.0, 0.0};
  int hv_timed_[2] = {0, 0};

 public:
  // pass 0: eigenvalue pass, 1: eigenvector pass
  double hv_ms(int pass = 0) const {
    return hv_timed_[pass] ? hv_ms_sum_[pass] / hv_timed_[pass] : 0.0;
  }
  int hv_timed(int pass = 0) const { return hv_timed_[pass]; }
  void release_events() {
    for (auto &e : ev_) {
      (void)hipEventDestroy(e.start);
      (void)hipEventDestroy(e.stop);
    }
    ev_.clear();
  }
};

nh_status LanczosRun::collect_timing() {
  NH_HIP(hipStreamSynchronize(c_->stream));
  for (auto &e : ev_) {
    float ms = 0.f;
    NH_HIP(hipEventElapsedTime(&ms, e.start, e.stop));
    const int pass = e.value_pass ? 0 : 1;
    hv_ms_sum_[pass] += ms;
    ++hv_timed_[pass];
  }
  release_events();
  return NH_OK;
}

__global__ __launch_bounds__(256) void k_to_f32(int64_t n,
                                                const double2 *__restrict__ x,
                                                float2 *__restrict__ y) {
  for (int64_t k = blockIdx.x * (int64_t)256 + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * 256)
    y[k] = make_float2((float)x[k].x, (float)x[k].y);
}

// start vector (lanczos_initialize, eigsys3d_par.c:113-215) built and
// normalised in device memory: CyclicRandom r0[j] = (s[2j mod L],
// s[2j+1 mod L]), Const1 (1, 1), or a caller's host or device vector
__global__ __launch_bounds__(256) void k_start_fill(int64_t n, int kind, int L,
                                                   const double *__restrict__ cyc,
                                                   double2 *__restrict__ r) {
  for (int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * 256)
    r[j] = kind == 0 ? make_double2(cyc[(2 * j) % L], cyc[(2 * j + 1) % L])
                     : make_double2(1.0, 1.0);
}
__global__ __launch_bounds__(256) void k_div(int64_t n, double2 *__restrict__ r,
                                             double s) {
  for (int64_t j = blockIdx.x * (int64_t)256 + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * 256)
    r[j] = make_double2(r[j].x / s, r[j].y / s);
}

nh_status LanczosRun::build_start(const nh_lanczos_opts *o) {
  NH_TRY(alloc((void **)&R0_, n_ * sizeof(double2)));
  const int g = reduce_blocks(n_);
  if (o->start == 0 || o->start == 1) {
    double *cyc = nullptr;
    int L = 1;
    if (o->start == 0) {
      NH_CHECK(o->cyclic_len > 0 && o->cyclic_string, NH_ERR_ARG,
               "CyclicRandom start needs CyclicString");
      L = o->cyclic_len;
      NH_TRY(alloc((void **)&cyc, (size_t)L * sizeof(double)));
      NH_HIP(hipMemcpyAsync(cyc, o->cyclic_string, (size_t)L * sizeof(double),
                            hipMemcpyHostToDevice, c_->stream));
    }
    hipLaunchKernelGGL(k_start_fill, dim3(g), dim3(256), 0, c_->stream, n_,
                       o->start, L, cyc, R0_);
    NH_HIP(hipGetLastError());
    // the host CyclicString may be released on return
    NH_HIP(hipStreamSynchronize(c_->stream));
  } else if (o->start == 4) {
    // Random: glibc rand() after srand(random_seed), drawn on the host in
    // the reference's order (one number per entry, real part)
    std::vector<nh_cplx> h((size_t)n_);
    rand_start(o->random_seed, n_, h.data());
    NH_HIP(hipMemcpyAsync(R0_, h.data(), n_ * sizeof(double2),
                          hipMemcpyHostToDevice, c_->stream));
    NH_HIP(hipStreamSynchronize(c_->stream));  // h is released on return
  } else {
    NH_CHECK(o->start == 2 || o->start == 3, NH_ERR_ARG, "bad start %d", o->start);
    NH_CHECK(o->start_vector, NH_ERR_ARG, "start vector missing");
    NH_HIP(hipMemcpyAsync(R0_, o->start_vector, n_ * sizeof(double2),
                          o->start == 2 ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                          c_->stream));
  }
  double s = 0.0;
  NH_TRY(reduce_nrm2sq(c_, n_, reinterpret_cast<const nh_cplx *>(R0_), &s));
  NH_CHECK(s > 0, NH_ERR_ARG, "zero start vector");
  hipLaunchKernelGGL(k_div, dim3(g), dim3(256), 0, c_->stream, n_, R0_, std::sqrt(s));
  NH_HIP(hipGetLastError());
  return NH_OK;
}

// r <- r0 (and its FP32 copy f32 when the products are FP32; only the
// latter with FP32 Lanczos vectors)
nh_status LanczosRun::upload_start(float2 *f32) {
  if (!v32_)
    NH_HIP(hipMemcpyAsync(R_, R0_, n_ * sizeof(double2), hipMemcpyDeviceToDevice,
                          c_->stream));
  if (f32) {
    hipLaunchKernelGGL(k_to_f32, dim3(reduce_blocks(n_)), dim3(256), 0,
                       c_->stream, n_, R0_, f32);
    NH_HIP(hipGetLastError());
  }
  return NH_OK;
}

nh_status LanczosRun::iterate(int k, bool red) {
  nh_context *c = c_;
  int sk_blocks = 0;
  const bool split = sk_like(c) && sk_split_lanes(c);
  const SkGrid sg = sk_grid(c->n_cations, c->n_anions, sk_mixed_roles(c), &sk_blocks,
                            split ? SkSplit::kLanes : 4);
  const int gridA = sk_like(c)
                        ? sk_blocks
                        : (c->n_atoms + kAtomsPerBlock - 1) / kAtomsPerBlock;
  // q_k is stored (ring slot) only in the eigenvector pass, which
  // accumulates from it; the update forms q_k, q_{k-1} from r_k, r_{k-1}
  double2 *Q = red || ring32_ || ringr_ ? nullptr : ring_ + (int64_t)(k % ring_k_) * n_;
  // FP64 vectors of the step: r_k, r_(k-1), r_(k+1)
  double2 *Rk = R_, *Rkm1 = Rm1_, *Yk = Y_;
  if (!red && ringr_) {
    Rk = ringr_ + (int64_t)(k % (2 * ring_k_)) * n_;
    Rkm1 = ringr_ + (int64_t)((k + 2 * ring_k_ - 1) % (2 * ring_k_)) * n_;
    Yk = ringr_ + (int64_t)((k + 1) % (2 * ring_k_)) * n_;
  }
  // FP32 copies: gathered (r_k) and written (r_{k+1})
  const float2 *g32 = R32_;
  float2 *y32 = Y32_;
  if (!red && ring32_) {
    g32 = ring32_ + (int64_t)(k % (2 * ring_k_)) * n_;
    y32 = ring32_ + (int64_t)((k + 1) % (2 * ring_k_)) * n_;
  }
  const double *bprev = k > 0 ? beta_ + (k - 1) : nullptr;
  const double *bprev2 = k > 1 ? beta_ + (k - 2) : nullptr;
  double *part = red ? part_ : nullptr;
  const bool timed = (k % kTimeEvery) == 0;
  // the eigenvector pass knows alpha_k: the SK kernel applies the update
  const bool fuse = !red && sk_like(c);
  // FP32 Lanczos vectors.  Eigenvalue pass: the product reads u_{k-1}
  // (own entries and gathers) and r_{k-1}, forms r_k, writes r_k and u_k
  // (step 0: u_{-1} = r_{-1} = r_0 with alpha_{-1} = beta_0 = 0, which
  // leaves r_0 as it is; nothing is written back for it).  Eigenvector
  // pass: own entries and gathers from ring slot k, r_{k-1} from slot k - 1
  // (slot 0 itself at step 0, the eigenvalue pass's operands), r_{k+1} to
  // slot k + 1; no FP64 vector at all.
  const float2 *rp32 = nullptr;
  float2 *rout = nullptr;
  const double *aprev = nullptr;
  if (v32_) {
    if (red) {
      g32 = k == 0 ? R32_ : U32_;
      rp32 = R32_;
      rout = k == 0 ? nullptr : Y32_;
      y32 = UN32_;
      aprev = k > 0 ? alpha_ + (k - 1) : nullptr;
      Yk = nullptr;
      part = part_;
    } else {
      rp32 = ring32_ + (int64_t)((k + 2 * ring_k_ - (k > 0 ? 1 : 0)) % (2 * ring_k_)) * n_;
      Yk = nullptr;
    }
  }
  const double2 *rp64 = k > 0 ? Rkm1 : nullptr;
  Timed evp{nullptr, nullptr, red};
  if (timed) {
    NH_HIP(hipEventCreate(&evp.start));
    NH_HIP(hipEventCreate(&evp.stop));
    NH_HIP(hipEventRecord(evp.start, c->stream));
  }
#define LZ_HV(NSP, TB, BLK)                                                    \
  do {                                                                         \
    if (use_staged(c))                                                          \
      hipLaunchKernelGGL((k_lz_hv<NSP, TB, TB, true>), dim3(gridA),            \
                         dim3(kThreads), 0, c->stream, c->n_atoms, c->onsite,  \
                         BLK, c->nbr, R_, Q, Y_, bprev, part);                 \
    else                                                                       \
      hipLaunchKernelGGL((k_lz_hv<NSP, TB, TB, false>), dim3(gridA),           \
                         dim3(kThreads), 0, c->stream, c->n_atoms, c->onsite,  \
                         BLK, c->nbr, R_, Q, Y_, bprev, part);                 \
  } while (0)
// XO, RV, RP: storage type of r, own entries r_k, r_{k-1}
#define LZ_SK2(NSP, R, SRC, QUAD, FUSE, XO, RV, RP)                             \
  hipLaunchKernelGGL((k_lz_hv_sk<NSP, R, QUAD, FUSE, decltype(SRC), XO>),       \
                     dim3(gridA), dim3(kSkThreads), 0, c->stream, c->n_cations, \
                     c->cations, c->n_anions, c->anions, sg, c->onsite, SRC,    \
                     c->nbr, RV, g32, Q, Yk, bprev, part, alpha_ + k, RP,       \
                     bprev2, y32, rout, aprev)
#define LZ_SK1(NSP, R, SRC, QUAD, XO, RV, RP)                                   \
  do {                                                                          \
    if (fuse) LZ_SK2(NSP, R, SRC, QUAD, true, XO, RV, RP);                      \
    else LZ_SK2(NSP, R, SRC, QUAD, false, XO, RV, RP);                          \
  } while (0)
#define LZ_SPL2(R, SRC, FUSE, XO, RV, RP)                                        \
  hipLaunchKernelGGL((k_lz_hv_sk_split<R, FUSE, decltype(SRC), XO>),            \
                     dim3(gridA), dim3(kSkThreads), 0, c->stream, c->n_cations, \
                     c->cations, c->n_anions, c->anions, sg, c->onsite, SRC,    \
                     c->nbr, RV, g32, Q, Yk, bprev, part, alpha_ + k, RP,       \
                     bprev2, y32, rout, aprev)
#define LZ_SK0(NSP, R, SRC, XO, RV, RP)                                         \
  do {                                                                          \
    if (NSP == 2 && split) {                                                    \
      if (fuse) LZ_SPL2(R, SRC, true, XO, RV, RP);                              \
      else LZ_SPL2(R, SRC, false, XO, RV, RP);                                  \
    } else if (use_staged(c)) LZ_SK1(NSP, R, SRC, true, XO, RV, RP);            \
    else LZ_SK1(NSP, R, SRC, false, XO, RV, RP);                                \
  } while (0)
#define LZ_SK(R, SRC, XO, RV, RP)                                               \
  do {                                                                          \
    if (c->nb == 20) LZ_SK0(2, R, SRC, XO, RV, RP);                             \
    else LZ_SK0(1, R, SRC, XO, RV, RP);                                         \
  } while (0)
// FP32 vectors (LDS gather only)
#define LZ_SKV0(NSP, SRC)                                                       \
  do {                                                                          \
    if (NSP == 2 && split) {                                                    \
      if (fuse) LZ_SPL2(float, SRC, true, float2, g32, rp32);                   \
      else LZ_SPL2(float, SRC, false, float2, g32, rp32);                       \
    } else LZ_SK1(NSP, float, SRC, true, float2, g32, rp32);                    \
  } while (0)
#define LZ_SKV(SRC)                                                             \
  do {                                                                          \
    if (c->nb == 20) LZ_SKV0(2, SRC);                                           \
    else LZ_SKV0(1, SRC);                                                       \
  } while (0)
  if (c->layout == NH_LAYOUT_MATRIX_FREE) {
    if (c->precision == NH_PREC_FP64) LZ_SK(double, mf_src<double>(c), double2, Rk, rp64);
    else if (v32_) LZ_SKV(mf_src<float>(c));
    else LZ_SK(float, mf_src<float>(c), double2, Rk, rp64);
  } else if (c->layout == NH_LAYOUT_SK) {
    if (c->precision == NH_PREC_FP64)
      LZ_SK(double, (SkSrc<double>{c->sk64, c->bond_of}), double2, Rk, rp64);
    else if (v32_) LZ_SKV((SkSrc<float>{c->sk32, c->bond_of}));
    else LZ_SK(float, (SkSrc<float>{c->sk32, c->bond_of}), double2, Rk, rp64);
  } else if (c->precision == NH_PREC_FP64) {
    if (c->nb == 20) LZ_HV(2, double, c->blocks64);
    else LZ_HV(1, double, c->blocks64);
  } else {
    if (c->nb == 20) LZ_HV(2, float, c->blocks32);
    else LZ_HV(1, float, c->blocks32);
  }
#undef LZ_HV
#undef LZ_SK
#undef LZ_SK0
#undef LZ_SK1
#undef LZ_SK2
#undef LZ_SPL2
#undef LZ_SKV
#undef LZ_SKV0
  if (timed) {
    NH_HIP(hipEventRecord(evp.stop, c->stream));
    ev_.push_back(evp);
  }
  if (v32_ && red) {
    // alpha_k, beta_{k+1} from the product's partials; rotate
    // (r_{k-1}, u_{k-1}) <- (r_k, u_k) (step 0: r_0 stays)
    hipLaunchKernelGGL(k_lz_scalars3, dim3(1), dim3(kRedThreads), 0, c->stream,
                       gridA, part_, alpha_ + k, beta_ + k);
    NH_HIP(hipGetLastError());
    if (k > 0) std::swap(R32_, Y32_);
    std::swap(U32_, UN32_);
    return NH_OK;
  }
  if (red)
    hipLaunchKernelGGL(k_lz_scalar, dim3(1), dim3(kRedThreads), 0, c->stream,
                       gridA, 1, part_, alpha_ + k, 0);
  const int nb = reduce_blocks(n_);
  if (!fuse)
    hipLaunchKernelGGL(k_lz_update, dim3(nb), dim3(kRedThreads), 0, c->stream,
                       n_, Y_, R_, k > 0 ? Rm1_ : nullptr, alpha_ + k, bprev,
                       bprev2, part, y32);
  if (red)
    hipLaunchKernelGGL(k_lz_scalar, dim3(1), dim3(kRedThreads), 0, c->stream,
                       nb, 2, part_, beta_ + k, 1);
  NH_HIP(hipGetLastError());
  // r_{k+1} is in Y: rotate (r_{k-1}, r_k, y) -> (r_k, r_{k+1}, free)
  if (v32_) {
    // eigenvector pass: the ring slots rotate by index
  } else {
    double2 *free_buf = Rm1_;
    Rm1_ = R_;
    R_ = Y_;
    Y_ = free_buf;
    std::swap(R32_, Y32_);
  }
  return NH_OK;
}

bool LanczosRun::window_done(int w, int req, int rule) const {
  if (req <= 0) return true;
  if (rule == 0) {
    // reference rule (lanczos.c:1231-1232): enough converged values found
    // anywhere in the window
    int cnt = 0;
    for (const auto &x : conv_) cnt += x.window == w;
    return cnt >= req;
  }
  // the `req` Ritz values nearest the gap (top of the vb window, bottom of
  // the cb window) must all be converged
  const auto &r = lastRitz_[w];
  const auto &cv = lastConv_[w];
  if ((int)r.size() < req) return false;
  for (int t = 0; t < req; ++t) {
    const int idx = w == 0 ? (int)r.size() - 1 - t : t;
    if (!cv[idx]) return false;
  }
  return true;
}

// queue the copy of alpha and beta up to step m into the pinned staging
nh_status LanczosRun::fetch(int m) {
  if (m > fetched_) {
    NH_HIP(hipMemcpyAsync(pin_ + fetched_, alpha_ + fetched_,
                          (m - fetched_) * sizeof(double), hipMemcpyDeviceToHost,
                          c_->stream));
    NH_HIP(hipMemcpyAsync(pin_ + pin_cap_ + fetched_, beta_ + fetched_,
                          (m - fetched_) * sizeof(double), hipMemcpyDeviceToHost,
                          c_->stream));
    fetched_ = m;
  }
  NH_HIP(hipEventRecord(fetch_ev_, c_->stream));
  return NH_OK;
}

// analyse the tridiagonal of the first m steps (fetch(m) queued before)
nh_status LanczosRun::analyse(int m, const nh_lanczos_opts *o) {
  NH_CHECK(fetched_ >= m, NH_ERR_STATE, "Lanczos: tridiagonal not fetched");
  NH_HIP(hipEventSynchronize(fetch_ev_));
  ha_.assign(pin_, pin_ + m);
  hb_.assign(pin_ + pin_cap_, pin_ + pin_cap_ + m);
  const double tol = o->conv_tolerance;
  const double resolution = 100.0 * tol;  // lanczos.c:1970
  const double bias = 1e-80;
  const int req[2] = {o->num_eig_vb, o->num_eig_cb};
  const double *win[2] = {o->erange_vb, o->erange_cb};
  // nearest-the-gap rule: two converged copies of one level differ by up to
  // about 2 tol ||T|| (each Ritz value within beta tol of it), which the
  // reference's relative resolution misses for levels near 0 eV; that rule
  // merges copies within this bound as well (the reference rule keeps the
  // reference's check)
  double dup_abs = 0.0;
  if (o->stop_rule == 1) {
    double tn = 0.0;
    for (int i = 0; i < m; ++i)
      tn = std::max(tn, std::fabs(ha_[i]) + (i > 0 ? std::fabs(hb_[i - 1]) : 0.0) +
                            (i + 1 < m ? std::fabs(hb_[i]) : 0.0));
    dup_abs = 2.0 * tol * tn;
  }
  for (int w = 0; w < 2; ++w) {
    lastRitz_[w].clear();
    lastConv_[w].clear();
    if (req[w] <= 0) continue;
    const double lo = win[w][0], hi = win[w][1];
    const int klo = sturm(ha_, hb_, m, lo), khi = sturm(ha_, hb_, m, hi);
    // candidates nearest the gap: top of vb, bottom of cb
    const int want = req[w] + 24;
    int k0 = klo, k1 = khi;
    if (w == 0) k0 = std::max(klo, khi - want);
    else k1 = std::min(khi, klo + want);
    // bisection of the candidates and the trailing components of their
    // tridiagonal eigenvectors, each independent: spread over host threads
    std::vector<double> ev(std::max(0, k1 - k0));
    host_parallel_for((int)ev.size(), [&](int j) {
      ev[j] = bisect(ha_, hb_, m, k0 + j, lo, hi, 1e-13);
    });
    auto duplicate = [&](size_t i) {
      // copies of one eigenvalue (ghosts, degeneracies) count once
      return i > 0 && std::fabs(ev[i - 1] - ev[i]) <
                          std::max(std::fabs(bias + ev[i]) * resolution, dup_abs);
    };
    auto known_value = [&](double v) {
      for (const auto &cv : conv_)
        if (std::fabs(cv.value - v) < resolution) return true;
      return false;
    };
    std::vector<double> tail((size_t)2 * ev.size(), 1.0);
    host_parallel_for((int)ev.size(), [&](int j) {
      if (duplicate(j) || known_value(ev[j])) return;
      std::vector<double> xj;
      tri_eigvec(ha_, hb_, m, ev[j], xj);
      tail[2 * j] = xj[m - 1];
      tail[2 * j + 1] = m >= 2 ? xj[m - 2] : 0.0;
    });
    for (size_t i = 0; i < ev.size(); ++i) {
      if (duplicate(i)) continue;
      const bool known = known_value(ev[i]);
      bool ok = known;
      if (!known) {
        ok = std::fabs(tail[2 * i]) <= tol && std::fabs(tail[2 * i + 1]) <= tol;
        if (ok) conv_.push_back({ev[i], m, w});
      }
      lastRitz_[w].push_back(ev[i]);
      lastConv_[w].push_back(ok);
    }
  }
  return NH_OK;
}

nh_status LanczosRun::run(const nh_lanczos_opts *o, nh_eig_result *res) {
  nh_context *c = c_;
  const auto t0 = std::chrono::steady_clock::now();
  NH_CHECK(c->layout == NH_LAYOUT_BLOCKS || sk_like(c), NH_ERR_STATE,
           "Lanczos: build the Hamiltonian first");
  NH_CHECK(o->max_iter > 2 && o->conv_tolerance > 0, NH_ERR_ARG,
           "Lanczos: bad max_iter / tolerance");
  n_ = c->n;
  v32_ = lz_fp32_vectors(c);
  if (!v32_) {
    NH_TRY(alloc((void **)&R_, n_ * sizeof(double2)));
    NH_TRY(alloc((void **)&Rm1_, n_ * sizeof(double2)));
  }
  NH_TRY(alloc((void **)&Y_, n_ * sizeof(double2)));
  NH_TRY(alloc((void **)&alpha_, (size_t)(o->max_iter + 1) * sizeof(double)));
  NH_TRY(alloc((void **)&beta_, (size_t)(o->max_iter + 1) * sizeof(double)));
  nparts_ = std::max((c->n_atoms + kAtomsPerBlock - 1) / kAtomsPerBlock,
                     2 * kRedMaxBlocks) + 64;
  NH_TRY(alloc((void **)&part_, (size_t)nparts_ * 3 * sizeof(double)));

  NH_TRY(build_start(o));
  if (sk_like(c) && c->precision == NH_PREC_FP32) {
    NH_TRY(alloc((void **)&R32_, n_ * sizeof(float2)));
    NH_TRY(alloc((void **)&Y32_, n_ * sizeof(float2)));
    if (v32_) {
      NH_TRY(alloc((void **)&U32_, n_ * sizeof(float2)));
      NH_TRY(alloc((void **)&UN32_, n_ * sizeof(float2)));
      int blocks = 0;
      sk_grid(c->n_cations, c->n_anions, sk_mixed_roles(c), &blocks,
              sk_split_lanes(c) ? SkSplit::kLanes : 4);
      NH_CHECK(blocks <= nparts_, NH_ERR_STATE, "Lanczos: partial sums do not fit");
    }
  }
  NH_TRY(upload_start(R32_));

  // pass 1: eigenvalues.  The host analyses the tridiagonal of one check
  // point while the device runs the steps up to the next one: at a check
  // point the new alpha, beta are copied behind an event, and they are
  // analysed at the following check point, after that block of steps has
  // been queued.  Convergence is thus seen one block late; the steps of
  // that block are not used (the run ends at the converged check point).
  pin_cap_ = o->max_iter + 1;
  NH_HIP(hipHostMalloc((void **)&pin_, 2 * (size_t)pin_cap_ * sizeof(double),
                       hipHostMallocDefault));
  NH_HIP(hipEventCreateWithFlags(&fetch_ev_, hipEventDisableTiming));
  int m = 0, skip = 0, pending = -1;
  bool done = false;
  int matvecs = 0;
  Progress prog;
  if (o->restore_len > 0) {
    // restart from the tridiagonal of an earlier run (restoreState,
    // lanczos.c:1905-1960): no eigenvalue pass
    NH_CHECK(o->restore_alpha && o->restore_beta && o->restore_len <= o->max_iter,
             NH_ERR_ARG, "Lanczos restart: tridiagonal missing or longer than max_iter");
    m = o->restore_len;
    NH_HIP(hipMemcpyAsync(alpha_, o->restore_alpha, (size_t)m * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
    NH_HIP(hipMemcpyAsync(beta_, o->restore_beta, (size_t)m * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
    NH_TRY(fetch(m));
    // the analyses of the original run, at the same check points (which
    // converged values were recorded depends on where they were seen)
    int mm = 0, sk = 0;
    while (mm < m) {
      ++mm;
      ++sk;
      if (o->conv_check_start < mm && o->conv_check_skip <= sk) {
        sk = 0;
        NH_TRY(analyse(mm, o));
        done = window_done(0, o->num_eig_vb, o->stop_rule) &&
               window_done(1, o->num_eig_cb, o->stop_rule);
        if (done) break;
      }
    }
    if (!done) {
      NH_TRY(analyse(m, o));
      done = window_done(0, o->num_eig_vb, o->stop_rule) &&
             window_done(1, o->num_eig_cb, o->stop_rule);
    } else {
      m = mm;
    }
    pending = m;
  }
  while (o->restore_len <= 0 && m + 1 < o->max_iter) {
    NH_TRY(iterate(m, true));
    ++m;
    prog.tick("lanczos", m, (int)conv_.size());
    ++matvecs;
    ++skip;
    if (o->conv_check_start < m && o->conv_check_skip <= skip) {
      skip = 0;
      if (pending > 0) {
        NH_TRY(analyse(pending, o));
        done = window_done(0, o->num_eig_vb, o->stop_rule) &&
               window_done(1, o->num_eig_cb, o->stop_rule);
        if (done) break;
      }
      NH_TRY(fetch(m));
      pending = m;
    }
  }
  if (done) {
    m = pending;
  } else if (o->restore_len <= 0) {
    NH_TRY(fetch(m));
    NH_TRY(analyse(m, o));
    done = window_done(0, o->num_eig_vb, o->stop_rule) &&
           window_done(1, o->num_eig_cb, o->stop_rule);
  }
  if (res->tridiag_alpha && res->tridiag_beta) {
    const int cnt = std::min(m, res->tridiag_cap);
    std::copy(ha_.begin(), ha_.begin() + cnt, res->tridiag_alpha);
    std::copy(hb_.begin(), hb_.begin() + cnt, res->tridiag_beta);
  }

  // the converged eigenvalues nearest the gap in each window
  std::vector<Converged> pick;
  const int req[2] = {o->num_eig_vb, o->num_eig_cb};
  for (int w = 0; w < 2; ++w) {
    if (o->stop_rule == 0) {
      // every converged value of the window, nearest the gap first
      std::vector<Converged> in;
      for (const auto &x : conv_)
        if (x.window == w) in.push_back(x);
      std::sort(in.begin(), in.end(), [w](const Converged &a, const Converged &b) {
        return w == 0 ? a.value > b.value : a.value < b.value;
      });
      for (int t = 0; t < (int)in.size() && t < req[w]; ++t) pick.push_back(in[t]);
      continue;
    }
    const auto &r = lastRitz_[w];
    const auto &cv = lastConv_[w];
    int got = 0;
    for (int t = 0; t < (int)r.size() && got < req[w]; ++t) {
      const int idx = w == 0 ? (int)r.size() - 1 - t : t;
      // report only the converged prefix nearest the gap: an unconverged
      // Ritz value there may be a state that a farther converged one would
      // wrongly replace
      if (!cv[idx]) break;
      for (const auto &x : conv_)
        if (std::fabs(x.value - r[idx]) < 100.0 * o->conv_tolerance) {
          pick.push_back({r[idx], x.iter, w});
          break;
        }
      ++got;
    }
  }
  std::sort(pick.begin(), pick.end(),
            [](const Converged &a, const Converged &b) { return a.value < b.value; });
  const int ne = std::min((int)pick.size(), std::min(res->capacity, kMaxVec));
  res->n_found = ne;
  res->iterations = m;
  for (int e = 0; e < ne; ++e) {
    res->eigenvalues[e] = pick[e].value;
    if (res->residuals) res->residuals[e] = -1.0;
  }

  if (o->want_vectors && ne > 0) {
    // pass 2: rebuild the Lanczos vectors from r0 and accumulate
    int mmax = 0;
    for (int e = 0; e < ne; ++e) mmax = std::max(mmax, pick[e].iter);
    std::vector<std::vector<double>> z(ne);
    for (int e = 0; e < ne; ++e)
      tri_eigvec(ha_, hb_, pick[e].iter, pick[e].value, z[e]);
    double2 *V = nullptr;
    double *dZ = nullptr;
    NH_TRY(alloc((void **)&V, (size_t)ne * n_ * sizeof(double2)));
    NH_TRY(alloc((void **)&dZ, (size_t)kMaxVec * kRingMax * sizeof(double)));
    NH_HIP(hipMemsetAsync(V, 0, (size_t)ne * n_ * sizeof(double2), c->stream));
    // ring: K = 128 steps per flush when the ring takes at most half the
    // free memory (the run's buffers are released before any later stage
    // allocates), else 32, 16 or 8 (each flush reads and writes the
    // eigenvectors once)
    const bool rring = !R32_ && sk_like(c);
    {
      size_t fr = 0, tot = 0;
      NH_HIP(hipMemGetInfo(&fr, &tot));
      fr += pool_free_bytes(c);
      const size_t slot = R32_ ? 2 * n_ * sizeof(float2)
                               : (rring ? 2 : 1) * n_ * sizeof(double2);
      ring_k_ = kRingMax;
      while (ring_k_ > 8 && (size_t)ring_k_ * slot > fr / 2) ring_k_ /= 2;
      if (c->ring_steps > 0) ring_k_ = std::min(ring_k_, c->ring_steps);
    }
    if (R32_) {
      NH_CHECK(n_ % 2 == 0, NH_ERR_STATE, "Lanczos: FP32 ring needs an even N");
      NH_TRY(alloc((void **)&ring32_, (size_t)2 * ring_k_ * n_ * sizeof(float2)));
      NH_TRY(upload_start(ring32_));
    } else if (rring) {
      NH_TRY(alloc((void **)&ringr_, (size_t)2 * ring_k_ * n_ * sizeof(double2)));
      NH_HIP(hipMemcpyAsync(ringr_, R0_, n_ * sizeof(double2),
                            hipMemcpyDeviceToDevice, c->stream));
    } else {
      NH_TRY(alloc((void **)&ring_, (size_t)ring_k_ * n_ * sizeof(double2)));
      NH_TRY(upload_start(nullptr));
    }
    std::vector<double> hz((size_t)kMaxVec * kRingMax);
    const int rk = ring_k_;
    for (int k = 0; k < mmax; ++k) {
      NH_TRY(iterate(k, false));
      ++matvecs;
      prog.tick("lanczos eigenvectors", k + 1, mmax);
      if ((k + 1) % rk == 0 || k + 1 == mmax) {
        const int k0 = k + 1 - ((k % rk) + 1);
        const int K = k + 1 - k0;
        for (int e = 0; e < ne; ++e)
          for (int t = 0; t < K; ++t) {
            const int j = k0 + t;
            // q_j = r_j / beta_{j-1} for the rings of r (FP32 copies or FP64)
            const double sc = (ring32_ || ringr_) && j > 0 ? 1.0 / hb_[j - 1] : 1.0;
            hz[e * K + t] = j < pick[e].iter ? z[e][j] * sc : 0.0;
          }
        NH_HIP(hipMemcpyAsync(dZ, hz.data(), (size_t)ne * K * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
        int64_t g = (n_ + 255) / 256;
        if (g > 8192) g = 8192;
        // ring slots k0..k hold q_k0..q_k in order since k0 % K == 0
        // (FP32 ring: slots k0 % (2 K) + t hold r_k0+t)
        if (ring32_)
          hipLaunchKernelGGL(k_lz_accum32, dim3((int)g), dim3(256), 0,
                             c->stream, n_, ne, K, dZ,
                             ring32_ + (int64_t)(k0 % (2 * rk)) * n_, n_, V, n_);
        else if (ringr_)
          hipLaunchKernelGGL((k_lz_accum<double2>), dim3((int)g), dim3(256), 0,
                             c->stream, n_, ne, K, dZ,
                             ringr_ + (int64_t)(k0 % (2 * rk)) * n_, n_, V, n_);
        else
          hipLaunchKernelGGL((k_lz_accum<double2>), dim3((int)g), dim3(256), 0,
                             c->stream, n_, ne, K, dZ, ring_, n_, V, n_);
        NH_HIP(hipGetLastError());
        // the host array hz is reused next chunk: wait for the copy
        NH_HIP(hipStreamSynchronize(c->stream));
      }
    }
    // normalise, Rayleigh quotient and residual in FP64
    for (int e = 0; e < ne; ++e) {
      const nh_cplx *v = reinterpret_cast<const nh_cplx *>(V + (int64_t)e * n_);
      double nrm2;
      NH_TRY(reduce_nrm2sq(c, n_, v, &nrm2));
      int64_t g = (n_ + 255) / 256;
      if (g > 8192) g = 8192;
      hipLaunchKernelGGL(k_scale, dim3((int)g), dim3(256), 0, c->stream, n_,
                         V + (int64_t)e * n_, 1.0 / std::sqrt(nrm2));
      NH_TRY(launch_hv(c, 0.0, v, reinterpret_cast<nh_cplx *>(Y_), 1, n_));
      ++matvecs;
      double rq[2];
      NH_TRY(reduce_dot(c, n_, v, reinterpret_cast<nh_cplx *>(Y_), rq));
      const nh_cplx a = {-rq[0], 0.0}, b = {1.0, 0.0};
      NH_TRY(launch_axpby(c, n_, a, v, b, reinterpret_cast<nh_cplx *>(Y_)));
      double r2;
      NH_TRY(reduce_nrm2sq(c, n_, reinterpret_cast<nh_cplx *>(Y_), &r2));
      res->eigenvalues[e] = rq[0];
      if (res->residuals) res->residuals[e] = std::sqrt(r2);
      if (res->eigenvectors)
        NH_HIP(hipMemcpy(res->eigenvectors + (int64_t)e * n_, v,
                         n_ * sizeof(nh_cplx), hipMemcpyDefault));  // host or device
    }
  }
  NH_HIP(hipStreamSynchronize(c->stream));
  res->matvecs = matvecs;
  NH_TRY(collect_timing());
  res->hv_ms = hv_ms(0);
  res->hv_timed = hv_timed(0);
  res->hv_ms_vec = hv_ms(1);
  res->hv_timed_vec = hv_timed(1);
  res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return done ? NH_OK : NH_ERR_NOCONV;
}

}  // namespace
}  // namespace nh

extern "C" nh_status nh_eigensolve_lanczos(nh_context *ctx,
                                           const nh_lanczos_opts *opts,
                                           nh_eig_result *res) {
  NH_CHECK(ctx && opts && res && res->eigenvalues && res->capacity > 0,
           NH_ERR_ARG, "nh_eigensolve_lanczos: bad arguments");
  (void)hipGetLastError();
  NH_HIP(hipSetDevice(ctx->device));
  nh::LanczosRun run(ctx);
  return run.run(opts, res);
}
