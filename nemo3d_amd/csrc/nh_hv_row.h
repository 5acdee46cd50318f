// Row of the stored-block H·v, shared by the plain operator kernel and
// the fused Lanczos kernels (the library is built without relocatable
// device code, so every kernel's translation unit includes this).
//
// Row-owner mapping: 10 lanes per atom, lane o owns spatial orbital o of
// both spins, so the spin-shared 10x10 real neighbour blocks
// (HamZincBlende::Hu, spin-diagonal and spin-independent) are read once per
// row and applied to both spin components (Hv_stored_20_Hzb,
// Hmult_stored.c:176-607, which applies Hu to rows i and i+10).  Each lane
// reads its block row as 10 contiguous values; the 10 lanes of an atom read
// one contiguous 800 B (FP64) / 400 B (FP32) block.
//
// Neighbour vector entries: either read by every lane straight from
// global memory (through L1/L2; hv_row), or first staged in LDS by the 10
// lanes of the atom, each fetching 8 of the 80 values once (stage_nbr_x +
// hv_row_staged).
//
// FP32 storage: each block product is formed in FP32 (values and the FP32
// image of x) and added to an FP64 row accumulator.
#pragma once

#include "nh_internal.h"

namespace nh {

constexpr int kAtomsPerBlock = 32;
constexpr int kThreads = kAtomsPerBlock * 10;  // 320 = 5 wavefronts

// Surface s-p coupling partner list for lane o in {1,2,3,4}: partner
// orbital and index into the 6 stored couplings (1,2),(1,3),(1,4),(2,3),
// (2,4),(3,4).
__device__ __forceinline__ void sp_partners(int o, int p[3], int k[3]) {
  switch (o) {
    case 1: p[0] = 2; k[0] = 0; p[1] = 3; k[1] = 1; p[2] = 4; k[2] = 2; break;
    case 2: p[0] = 1; k[0] = 0; p[1] = 3; k[1] = 3; p[2] = 4; k[2] = 4; break;
    case 3: p[0] = 1; k[0] = 1; p[1] = 2; k[1] = 3; p[2] = 4; k[2] = 5; break;
    default: p[0] = 1; k[0] = 2; p[1] = 2; k[1] = 4; p[2] = 3; k[2] = 5; break;
  }
}

// Onsite part of row (atom, spin s, orbital o): (diagonal - shift) x,
// dangling-bond s-p couplings and spin-orbit.  Initialises acc.
template <int NSP>
__device__ __forceinline__ void hv_onsite(int atom, int o,
                                          const double *__restrict__ onsite,
                                          double shift,
                                          const double2 *__restrict__ xc,
                                          double acc_re[NSP], double acc_im[NSP],
                                          double2 xi[NSP]) {
  constexpr int NB = NSP * 10;
  const int64_t base = (int64_t)atom * NB;
  const double *os = onsite + (size_t)atom * kOnsiteStride;
  const double d = os[o] - shift;
#pragma unroll
  for (int s = 0; s < NSP; ++s) {
    xi[s] = xc[base + s * 10 + o];
    acc_re[s] = d * xi[s].x;
    acc_im[s] = d * xi[s].y;
  }
  // dangling-bond s-p couplings (zero away from the surface)
  if (o >= 1 && o <= 4) {
    int p[3], k[3];
    sp_partners(o, p, k);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double h = os[10 + k[q]];
      if (h != 0.0) {
#pragma unroll
        for (int s = 0; s < NSP; ++s) {
          const double2 xp = xc[base + s * 10 + p[q]];
          acc_re[s] += h * xp.x;
          acc_im[s] += h * xp.y;
        }
      }
    }
  }
  // spin-orbit (h3d_diag + mb_ham_spds_so.h; pairs HamZincBlende.c:170-183)
  if (NSP == 2 && o >= 2 && o <= 4) {
    const double lam = os[16];
    const double2 *xa = xc + base;
    // up row o and down row 10+o
    if (o == 2) {        // (2,3) = -i lam, (2,14) = lam ; (12,13) = i lam, (12,4) = -lam
      acc_re[0] += lam * xa[3].y + lam * xa[14].x;
      acc_im[0] += -lam * xa[3].x + lam * xa[14].y;
      acc_re[NSP - 1] += -lam * xa[13].y - lam * xa[4].x;
      acc_im[NSP - 1] += lam * xa[13].x - lam * xa[4].y;
    } else if (o == 3) { // (3,2) = i lam, (3,14) = -i lam ; (13,12) = -i lam, (13,4) = -i lam
      acc_re[0] += -lam * xa[2].y + lam * xa[14].y;
      acc_im[0] += lam * xa[2].x - lam * xa[14].x;
      acc_re[NSP - 1] += lam * xa[12].y + lam * xa[4].y;
      acc_im[NSP - 1] += -lam * xa[12].x - lam * xa[4].x;
    } else {             // (4,12) = -lam, (4,13) = i lam ; (14,2) = lam, (14,3) = i lam
      acc_re[0] += -lam * xa[12].x - lam * xa[13].y;
      acc_im[0] += -lam * xa[12].y + lam * xa[13].x;
      acc_re[NSP - 1] += lam * xa[2].x - lam * xa[3].y;
      acc_im[NSP - 1] += lam * xa[2].y + lam * xa[3].x;
    }
  }
}

// One neighbour block row applied to both spins: acc += B[o][:] x_j.
template <int NSP, typename TB, typename TA>
__device__ __forceinline__ void hv_block(const TB *__restrict__ B,
                                         const double2 *xj, double acc_re[NSP],
                                         double acc_im[NSP]) {
  TA br[NSP], bi[NSP];
#pragma unroll
  for (int s = 0; s < NSP; ++s) br[s] = bi[s] = TA(0);
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const TA b = TA(B[k]);
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
      const double2 v = xj[s * 10 + k];
      br[s] += b * TA(v.x);
      bi[s] += b * TA(v.y);
    }
  }
#pragma unroll
  for (int s = 0; s < NSP; ++s) {
    acc_re[s] += double(br[s]);
    acc_im[s] += double(bi[s]);
  }
}

// acc[s] = ((H - shift I) x) at row (atom, spin s, orbital o).
template <int NSP, typename TB, typename TA>
__device__ __forceinline__ void hv_row(int atom, int o,
                                       const double *__restrict__ onsite,
                                       const TB *__restrict__ blocks,
                                       const int32_t *__restrict__ nbr,
                                       double shift,
                                       const double2 *__restrict__ xc,
                                       double acc_re[NSP], double acc_im[NSP],
                                       double2 xi[NSP]) {
  constexpr int NB = NSP * 10;
  hv_onsite<NSP>(atom, o, onsite, shift, xc, acc_re, acc_im, xi);
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int j = nbr[atom * 4 + n];
    if (j < 0) continue;
    hv_block<NSP, TB, TA>(blocks + ((size_t)atom * 4 + n) * 100 + o * 10,
                          xc + (int64_t)j * NB, acc_re, acc_im);
  }
}

// Stages the x entries of the atom's four neighbours (4 x NSP*10 complex
// values) in shared memory: lane o fetches entries o, o+10, ...; missing
// neighbours are staged as zeros.  sx: the atom's 40*NSP slots.
template <int NSP>
__device__ __forceinline__ void stage_nbr_x(int atom, int o,
                                            const int32_t *__restrict__ nbr,
                                            const double2 *__restrict__ xc,
                                            double2 *sx) {
  constexpr int NB = NSP * 10;
#pragma unroll
  for (int t = 0; t < 4 * NSP; ++t) {
    const int e = o + 10 * t;  // 0 .. 40*NSP-1
    const int n = e / NB, k = e % NB;
    const int j = nbr[atom * 4 + n];
    sx[e] = j >= 0 ? xc[(int64_t)j * NB + k] : make_double2(0.0, 0.0);
  }
}

// hv_row with the neighbour entries read from the staged copy.
template <int NSP, typename TB, typename TA>
__device__ __forceinline__ void hv_row_staged(int atom, int o,
                                              const double *__restrict__ onsite,
                                              const TB *__restrict__ blocks,
                                              double shift,
                                              const double2 *__restrict__ xc,
                                              const double2 *sx,
                                              double acc_re[NSP],
                                              double acc_im[NSP],
                                              double2 xi[NSP]) {
  constexpr int NB = NSP * 10;
  hv_onsite<NSP>(atom, o, onsite, shift, xc, acc_re, acc_im, xi);
#pragma unroll
  for (int n = 0; n < 4; ++n)
    hv_block<NSP, TB, TA>(blocks + ((size_t)atom * 4 + n) * 100 + o * 10,
                          sx + n * NB, acc_re, acc_im);
}

}  // namespace nh
