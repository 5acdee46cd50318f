// Sparse H·v and block H·V for the stored-block layout (row computation
// in nh_hv_row.h).
#include "nh_hv_row.h"
#include "nh_hv_sk.h"

namespace nh {
namespace {

// SK-compressed and matrix-free layouts (see nh_hv_sk.h; SRC = SkSrc or
// MfSrc); workgroup -> atoms by sk_lane_of.
template <int NSP, typename R, bool QUAD, typename SRC>
__global__ __launch_bounds__(kSkThreads) void k_hv_sk(
    int nc, const int32_t *__restrict__ cations, int na,
    const int32_t *__restrict__ anions, SkGrid grid,
    const double *__restrict__ onsite, const SRC src,
    const int32_t *__restrict__ nbr, double shift,
    const double2 *__restrict__ x, double2 *__restrict__ y, int64_t ld) {
  constexpr int NB = NSP * 10;
  const SkLane ln = sk_lane_of(grid, blockIdx.x, threadIdx.x);
  // the role is uniform across the wavefront; read as a scalar, the role
  // branches below are scalar branches (one per wave)
  const bool cat = __builtin_amdgcn_readfirstlane((int)ln.cat) != 0;
  const int li = ln.li;
  const int slot = threadIdx.x & 3;
  const bool active = li < (cat ? nc : na);
  const int atom = active ? (cat ? cations[li] : anions[li]) : 0;
  const double2 *xc = x + (int64_t)blockIdx.y * ld;
  double2 *yc = y + (int64_t)blockIdx.y * ld;
  constexpr int NO = (NB + 3) / 4;
  double own_re[NO], own_im[NO];
  SkQuadStage<NSP, double2> *L = nullptr;
  if constexpr (QUAD) {
    __shared__ SkQuadStage<NSP, double2> stage;
    L = &stage;
  }
  if (cat)
    sk_bonds<NSP, R, true, double2, QUAD>(atom, slot, active, src, nbr, xc,
                                          own_re, own_im, L);
  else
    sk_bonds<NSP, R, false, double2, QUAD>(atom, slot, active, src, nbr, xc,
                                           own_re, own_im, L);
  if constexpr (QUAD) {
    auto &S = *reinterpret_cast<SkOwnStage<NSP> *>(L->buf);
    const int la = threadIdx.x / 4;
    // own entries and onsite record loaded before the barrier (the loads
    // overlap it), staged in the freed LDS after it
    SkOwnRegs<NSP> own;
    if (active)
      sk_load_own<NSP>(slot, xc + (int64_t)atom * NB,
                       onsite + (size_t)atom * kOnsiteStride, own);
    __syncthreads();  // every lane has read its staged neighbour entries
    if (active) sk_store_own<NSP>(la, slot, own, S);
    __syncthreads();
    if (!active) return;
#pragma unroll
    for (int i = 0; i < NO; ++i) {
      const int r = 4 * i + slot;
      if (r >= NB) break;
      double re = own_re[i], im = own_im[i];
      onsite_row<NSP>(reinterpret_cast<const double *>(S.os[la]), S.x[la],
                      r / 10, r % 10, shift, re, im);
      yc[(int64_t)atom * NB + r] = make_double2(re, im);
    }
    return;
  }
  if (!active) return;
  const double *os = onsite + (size_t)atom * kOnsiteStride;
  const double2 *xa = xc + (int64_t)atom * NB;
  // all four lanes of the quad active on every step (row r = 4 i + slot)
#pragma unroll
  for (int i = 0; i < NO; ++i) {
    const int r = 4 * i + slot;
    if (r >= NB) break;
    double re = own_re[i], im = own_im[i];
    onsite_row<NSP>(os, xa, r / 10, r % 10, shift, re, im);
    yc[(int64_t)atom * NB + r] = make_double2(re, im);
  }
}

// Block H·V in the SK-compressed and matrix-free layouts: the columns of
// the block one after another inside each workgroup, the bond records,
// neighbour indices and onsite records read once for all of them
// (Hmatmult_spds_col_stored, Hmult_stored.c:2359-2474, re-reads H for every
// column).  Mapping and the per-column steps as k_hv_sk with the quad LDS
// gather.
template <int NSP, typename R, bool CATION, typename SRC>
__device__ __forceinline__ void sk_columns(int atom, int slot, bool active,
                                           const SRC &src,
                                           const int32_t *__restrict__ nbr,
                                           const double *__restrict__ onsite,
                                           double shift, const double2 *__restrict__ X,
                                           double2 *__restrict__ Y, int64_t ld,
                                           int ncol, SkQuadStage<NSP, double2> &L) {
  constexpr int NB = NSP * 10;
  constexpr int NO = (NB + 3) / 4;
  const int j = active ? nbr[atom * 4 + slot] : -1;
  R rv[24];
  if (j >= 0) src.template fetch<CATION>(atom, slot, j, rv);
  SkOwnRegs<NSP> own;
  auto &S = *reinterpret_cast<SkOwnStage<NSP> *>(L.buf);
  const int la = threadIdx.x / 4;
  for (int col = 0; col < ncol; ++col) {
    const double2 *xc = X + (int64_t)col * ld;
    double2 *yc = Y + (int64_t)col * ld;
    double own_re[NO], own_im[NO];
    sk_bonds_rec<NSP, R, CATION, double2>(j, slot, rv, xc, own_re, own_im, L);
    if (active)
      sk_load_own<NSP>(slot, xc + (int64_t)atom * NB,
                       onsite + (size_t)atom * kOnsiteStride, own);
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
                        r / 10, r % 10, shift, re, im);
        yc[(int64_t)atom * NB + r] = make_double2(re, im);
      }
    }
    __syncthreads();  // the epilogue has read the staging before it is reused
  }
}

template <int NSP, typename R, typename SRC>
__global__ __launch_bounds__(kSkThreads) void k_hV_sk(
    int nc, const int32_t *__restrict__ cations, int na,
    const int32_t *__restrict__ anions, SkGrid grid,
    const double *__restrict__ onsite, const SRC src,
    const int32_t *__restrict__ nbr, double shift,
    const double2 *__restrict__ X, double2 *__restrict__ Y, int64_t ld, int ncol) {
  // one role per workgroup (grid.mixed == 0): the role branch below holds
  // the column loop's barriers
  const SkLane ln = sk_lane_of(grid, blockIdx.x, threadIdx.x);
  // the role is uniform across the wavefront; read as a scalar, the role
  // branches below are scalar branches (one per wave)
  const bool cat = __builtin_amdgcn_readfirstlane((int)ln.cat) != 0;
  const int li = ln.li;
  const int slot = threadIdx.x & 3;
  const bool active = li < (cat ? nc : na);
  const int atom = active ? (cat ? cations[li] : anions[li]) : 0;
  __shared__ SkQuadStage<NSP, double2> stage;
  if (cat)
    sk_columns<NSP, R, true>(atom, slot, active, src, nbr, onsite, shift, X, Y,
                             ld, ncol, stage);
  else
    sk_columns<NSP, R, false>(atom, slot, active, src, nbr, onsite, shift, X, Y,
                              ld, ncol, stage);
}

template <int NSP, typename TB, typename TA, bool STAGED>
__global__ __launch_bounds__(kThreads) void k_hv_blocks(
    int A, const double *__restrict__ onsite, const TB *__restrict__ blocks,
    const int32_t *__restrict__ nbr, double shift,
    const double2 *__restrict__ x, double2 *__restrict__ y, int64_t ld) {
  constexpr int NB = NSP * 10;
  const int la = threadIdx.x / 10;
  const int atom = blockIdx.x * kAtomsPerBlock + la;
  const int o = threadIdx.x % 10;
  const double2 *xc = x + (int64_t)blockIdx.y * ld;
  double2 *yc = y + (int64_t)blockIdx.y * ld;
  double acc_re[NSP], acc_im[NSP];
  double2 xi[NSP];
  if constexpr (STAGED) {
    __shared__ double2 sx[kAtomsPerBlock * 4 * NB];
    if (atom < A) stage_nbr_x<NSP>(atom, o, nbr, xc, sx + la * 4 * NB);
    __syncthreads();
    if (atom >= A) return;
    hv_row_staged<NSP, TB, TA>(atom, o, onsite, blocks, shift, xc,
                               sx + la * 4 * NB, acc_re, acc_im, xi);
  } else {
    if (atom >= A) return;
    hv_row<NSP, TB, TA>(atom, o, onsite, blocks, nbr, shift, xc, acc_re,
                        acc_im, xi);
  }
  const int64_t base = (int64_t)atom * NB;
#pragma unroll
  for (int s = 0; s < NSP; ++s)
    yc[base + s * 10 + o] = make_double2(acc_re[s], acc_im[s]);
}

}  // namespace

nh_status launch_hv(nh_context *c, double shift, const nh_cplx *x, nh_cplx *y,
                    int ncol, int64_t ld) {
  NH_CHECK(c->layout == NH_LAYOUT_BLOCKS || sk_like(c), NH_ERR_STATE,
           "H·v: Hamiltonian not built in a supported layout");
  NH_CHECK(ncol >= 1 && ncol <= 65535 && ld >= c->n, NH_ERR_ARG,
           "H·V: bad ncol %d / ld %lld", ncol, (long long)ld);
  // block products: the matrix-core kernel (FP32, 8 columns), all the
  // columns in one SK / matrix-free launch with the bond records read once,
  // or one column per grid row.  Default (-1): the matrix-core kernel for
  // the row-owner block layout; one column per grid row for the SK and
  // matrix-free layouts, measured fastest on a 786,520-atom dot (8 columns:
  // 2.46 ms FP64 / 2.12 ms FP32 against 4.38 / 2.53 ms in one launch and
  // 4.02 ms on the matrix cores)
  const int mode = c->block_hv;
  if ((mode == 2 || (mode == -1 && c->layout == NH_LAYOUT_BLOCKS)) &&
      hV_mfma_supported(c, ncol))
    return launch_hV_mfma(c, shift, x, y, ld);
  if (sk_like(c) && ncol > 1 && mode == 1) {
    int blocks = 0;
    const SkGrid sg = sk_grid(c->n_cations, c->n_anions, false, &blocks);
    const double2 *xd = reinterpret_cast<const double2 *>(x);
    double2 *yd = reinterpret_cast<double2 *>(y);
#define HVK(NSP, R, SRC)                                                      \
  hipLaunchKernelGGL((k_hV_sk<NSP, R, decltype(SRC)>), dim3(blocks),          \
                     dim3(kSkThreads), 0, c->stream, c->n_cations,            \
                     c->cations, c->n_anions, c->anions, sg, c->onsite, SRC,  \
                     c->nbr, shift, xd, yd, ld, ncol)
#define HVP(R, SRC)                                                          \
  do {                                                                       \
    if (c->nb == 20) HVK(2, R, SRC);                                         \
    else HVK(1, R, SRC);                                                     \
  } while (0)
    if (c->layout == NH_LAYOUT_MATRIX_FREE) {
      if (c->precision == NH_PREC_FP64) HVP(double, mf_src<double>(c));
      else HVP(float, mf_src<float>(c));
    } else if (c->precision == NH_PREC_FP64) {
      HVP(double, (SkSrc<double>{c->sk64, c->bond_of}));
    } else {
      HVP(float, (SkSrc<float>{c->sk32, c->bond_of}));
    }
#undef HVP
#undef HVK
    NH_HIP(hipGetLastError());
    return NH_OK;
  }
  if (sk_like(c)) {
    int blocks = 0;
    const SkGrid sg = sk_grid(c->n_cations, c->n_anions, sk_mixed_roles(c), &blocks);
    const dim3 g(blocks, ncol);
    const double2 *xd = reinterpret_cast<const double2 *>(x);
    double2 *yd = reinterpret_cast<double2 *>(y);
#define SKL1(NSP, R, SRC, QUAD)                                               \
  hipLaunchKernelGGL((k_hv_sk<NSP, R, QUAD, decltype(SRC)>), g,               \
                     dim3(kSkThreads), 0, c->stream, c->n_cations,            \
                     c->cations, c->n_anions, c->anions, sg, c->onsite, SRC,  \
                     c->nbr, shift, xd, yd, ld)
#define SKL(NSP, R, SRC)                                                     \
  do {                                                                       \
    if (use_staged(c)) SKL1(NSP, R, SRC, true);                               \
    else SKL1(NSP, R, SRC, false);                                           \
  } while (0)
#define SKP(R, SRC)                                                          \
  do {                                                                       \
    if (c->nb == 20) SKL(2, R, SRC);                                         \
    else SKL(1, R, SRC);                                                     \
  } while (0)
    if (c->layout == NH_LAYOUT_MATRIX_FREE) {
      if (c->precision == NH_PREC_FP64) SKP(double, mf_src<double>(c));
      else SKP(float, mf_src<float>(c));
    } else if (c->precision == NH_PREC_FP64) {
      SKP(double, (SkSrc<double>{c->sk64, c->bond_of}));
    } else {
      SKP(float, (SkSrc<float>{c->sk32, c->bond_of}));
    }
#undef SKP
#undef SKL
#undef SKL1
    NH_HIP(hipGetLastError());
    return NH_OK;
  }
  const dim3 grid((c->n_atoms + kAtomsPerBlock - 1) / kAtomsPerBlock, ncol);
  const double2 *xd = reinterpret_cast<const double2 *>(x);
  double2 *yd = reinterpret_cast<double2 *>(y);
  const bool st = use_staged(c);
#define HVB(NSP, TB, TA, BLK)                                                  \
  do {                                                                       \
    if (st)                                                                  \
      hipLaunchKernelGGL((k_hv_blocks<NSP, TB, TA, true>), grid,             \
                         dim3(kThreads), 0, c->stream, c->n_atoms, c->onsite, \
                         BLK, c->nbr, shift, xd, yd, ld);                    \
    else                                                                     \
      hipLaunchKernelGGL((k_hv_blocks<NSP, TB, TA, false>), grid,            \
                         dim3(kThreads), 0, c->stream, c->n_atoms, c->onsite, \
                         BLK, c->nbr, shift, xd, yd, ld);                    \
  } while (0)
  if (c->precision == NH_PREC_FP64) {
    if (c->nb == 20) HVB(2, double, double, c->blocks64);
    else HVB(1, double, double, c->blocks64);
  } else {
    if (c->nb == 20) HVB(2, float, float, c->blocks32);
    else HVB(1, float, float, c->blocks32);
  }
#undef HVB
  NH_HIP(hipGetLastError());
  return NH_OK;
}

}  // namespace nh
