// Block H·V on the matrix cores: Y = (H - shift) X for a block of 8 column
// vectors, FP32 neighbour blocks and FP32 products with FP32 accumulation,
// FP64 onsite terms and FP64 output (the block-Lanczos / TRACEMIN product,
// Hmatmult_spds_col_stored, Hmult_stored.c:2359-2474, which loops the
// single-vector product over the columns and re-reads H for each one).
//
// One wavefront per atom.  The atom's four 10x10 neighbour blocks form a
// 10x40 matrix A (rows: the atom's orbitals; columns: neighbour n, orbital
// k); the neighbours' vector entries form a 40x(16*NSP) matrix B (columns:
// spin, vector, real/imaginary part).  C = A B runs as 16x16x4 FP32 MFMAs
// (rows padded 10 -> 16), so each H entry is read once for all 8 vectors
// instead of once per vector.  A comes from the stored FP32 blocks or, in the
// SK-compressed layout, is rebuilt in LDS from the bond records (four lanes,
// one bond each).  The epilogue adds the onsite part (diagonal - shift,
// surface s-p couplings, spin-orbit) in FP64 and writes Y.
#include "nh_hv_sk.h"

namespace nh {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kMfmaCols = 8;          // vectors per product
constexpr int kMfmaWaves = 4;         // atoms (wavefronts) per workgroup
constexpr int kMfmaThreads = kMfmaWaves * 64;

template <int NSP>
struct MfmaShape {
  static constexpr int NT = NSP;              // 16-column tiles of B / C
  static constexpr int BS = NT * 16 + 8;       // B/C row stride (floats)
  static constexpr int A_F = 400;             // A: [n][i][k]
  static constexpr int B_F = 40 * BS;         // B: [n*10+k][col]
  static constexpr int C_F = 10 * BS;         // C: [i][col], rows 0-9
  static constexpr int WAVE_F = A_F + B_F + C_F;
};

// Rebuilds T (anion-row orientation, T[a][c]) of one bond into dst[a*10+c]
// (anion role) or dst[c*10+a] (cation role).
template <typename R>
__device__ __forceinline__ void sk_block_to_lds(const R *__restrict__ rec,
                                                bool cation, float *dst) {
  R v[21];
#pragma unroll
  for (int k = 0; k < 21; ++k) v[k] = rec[3 + k];
  const R l = rec[0], m = rec[1], n = rec[2];
  R row[10];
#define NH_ROW(A)                                                   \
  sk_row<A, R>(l, m, n, v, row);                                    \
  _Pragma("unroll") for (int c = 0; c < 10; ++c)                    \
      dst[cation ? c * 10 + A : A * 10 + c] = float(row[c]);
  NH_ROW(0) NH_ROW(1) NH_ROW(2) NH_ROW(3) NH_ROW(4)
  NH_ROW(5) NH_ROW(6) NH_ROW(7) NH_ROW(8) NH_ROW(9)
#undef NH_ROW
}

template <int NSP, bool SK>
__global__ __launch_bounds__(kMfmaThreads) void k_hV_mfma(
    int A, const double *__restrict__ onsite, const float *__restrict__ blocks,
    const float *__restrict__ sk, const int32_t *__restrict__ bond_of,
    const int32_t *__restrict__ is_cation, const int32_t *__restrict__ nbr,
    double shift, const double2 *__restrict__ X, double2 *__restrict__ Y,
    int64_t ld, int nwg) {
  using S = MfmaShape<NSP>;
  constexpr int NB = NSP * 10;
  extern __shared__ float lds[];
  // consecutive workgroups go to different XCDs; give each XCD a contiguous
  // range of atoms so that neighbour vector entries are shared in its L2
  const int b = blockIdx.x;
  const int per = (nwg + 7) / 8;
  const int wg = (b % 8) * per + b / 8;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int atom = wg * kMfmaWaves + wave;
  const bool active = wg < nwg && atom < A;
  float *sA = lds + wave * S::WAVE_F;
  float *sB = sA + S::A_F;
  float *sC = sB + S::B_F;

  if (active) {
    // B: neighbour vector entries, FP32 (zero for a missing neighbour);
    // issued first so that the loads are in flight while A is rebuilt
    const int c = lane >> 3, t = lane & 7;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int j = nbr[atom * 4 + n];
      const double2 *xj = X + (int64_t)c * ld + (int64_t)(j < 0 ? 0 : j) * NB;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const int e = t + 8 * q;
        if (e < NB) {
          double2 v = j >= 0 ? xj[e] : make_double2(0.0, 0.0);
          const int s = e / 10, k = e % 10;
          *reinterpret_cast<float2 *>(sB + (n * 10 + k) * S::BS + (s * 8 + c) * 2) =
              make_float2(float(v.x), float(v.y));
        }
      }
    }
    // A: the four neighbour blocks
    if constexpr (!SK) {
      const float4 *src = reinterpret_cast<const float4 *>(blocks + (size_t)atom * 400);
      float4 *dst = reinterpret_cast<float4 *>(sA);
      dst[lane] = src[lane];
      if (lane < 100 - 64) dst[64 + lane] = src[64 + lane];
    } else {
      if (lane < 4) {
        const int j = nbr[atom * 4 + lane];
        float *dst = sA + lane * 100;
        if (j >= 0) {
          sk_block_to_lds<float>(sk + (size_t)bond_of[atom * 4 + lane] * kSkStride,
                                 is_cation[atom] != 0, dst);
        } else {
#pragma unroll
          for (int q = 0; q < 100; ++q) dst[q] = 0.f;
        }
      }
    }
  }
  __syncthreads();
  if (active) {
    const int i = lane & 15, kq = lane >> 4;
    float a[10];
#pragma unroll
    for (int ks = 0; ks < 10; ++ks) {
      const int K = ks * 4 + kq;
      a[ks] = i < 10 ? sA[(K / 10) * 100 + i * 10 + (K % 10)] : 0.f;
    }
#pragma unroll
    for (int tile = 0; tile < S::NT; ++tile) {
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < 10; ++ks) {
        const float bv = sB[(ks * 4 + kq) * S::BS + tile * 16 + i];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ks], bv, acc, 0, 0, 0);
      }
      // acc[r] = C[4*kq + r][tile*16 + i]; rows 10-15 are padding
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (4 * kq + r < 10) sC[(4 * kq + r) * S::BS + tile * 16 + i] = acc[r];
    }
  }
  __syncthreads();
  if (active) {
    const double *os = onsite + (size_t)atom * kOnsiteStride;
    for (int q = lane; q < kMfmaCols * NB; q += 64) {
      const int c = q / NB, r = q % NB, s = r / 10, o = r % 10;
      const double2 *xa = X + (int64_t)c * ld + (int64_t)atom * NB;
      const float2 cv =
          *reinterpret_cast<const float2 *>(sC + o * S::BS + (s * 8 + c) * 2);
      double re = double(cv.x), im = double(cv.y);
      onsite_row<NSP>(os, xa, s, o, shift, re, im);
      Y[(int64_t)c * ld + (int64_t)atom * NB + r] = make_double2(re, im);
    }
  }
}

}  // namespace

bool hV_mfma_supported(const nh_context *c, int ncol) {
  return ncol == kMfmaCols && c->precision == NH_PREC_FP32 &&
         (c->layout == NH_LAYOUT_BLOCKS || c->layout == NH_LAYOUT_SK);
}

nh_status launch_hV_mfma(nh_context *c, double shift, const nh_cplx *x,
                         nh_cplx *y, int64_t ld) {
  NH_CHECK(hV_mfma_supported(c, kMfmaCols), NH_ERR_STATE,
           "block H·V on MFMA needs an FP32 build");
  NH_CHECK(ld >= c->n, NH_ERR_ARG, "H·V: ld %lld < N", (long long)ld);
  const int nwg = (c->n_atoms + kMfmaWaves - 1) / kMfmaWaves;
  const int grid = ((nwg + 7) / 8) * 8;
  const double2 *xd = reinterpret_cast<const double2 *>(x);
  double2 *yd = reinterpret_cast<double2 *>(y);
  const bool sk = c->layout == NH_LAYOUT_SK;
#define NH_MFMA(NSP, SKL)                                                     \
  hipLaunchKernelGGL((k_hV_mfma<NSP, SKL>), dim3(grid), dim3(kMfmaThreads),   \
                     kMfmaWaves * MfmaShape<NSP>::WAVE_F * sizeof(float),     \
                     c->stream, c->n_atoms, c->onsite, c->blocks32, c->sk32,  \
                     c->bond_of, c->is_cation, c->nbr, shift, xd, yd, ld, nwg)
  if (c->nb == 20) {
    if (sk) NH_MFMA(2, true); else NH_MFMA(2, false);
  } else {
    if (sk) NH_MFMA(1, true); else NH_MFMA(1, false);
  }
#undef NH_MFMA
  NH_HIP(hipGetLastError());
  return NH_OK;
}

}  // namespace nh
