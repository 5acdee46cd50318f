// On-device assembly of the sp3d5s* Hamiltonian.
//
// k_bond: one thread per cation-owned bond (every zincblende bond joins one
// cation-role and one anion-role atom, so each bond is computed exactly
// once, as the reference stores each bond once in HamZincBlende::Hu).  It
// evaluates the strain-scaled Slater-Koster block and the Boykin diagonal
// corrections of both atoms (H_DiagStrainCorr_OffDiag_Wrapper +
// h3d_offdiag_diag_corr, h_cvectr_mult.c:1007-1126,
// mb_ham_spds_micro.c:597-1451).
// k_onsite: one thread per atom builds the onsite record: neighbour-averaged
// parameters, strain reference shift, dangling-bond passivation, potential
// and the summed corrections (H_Diag_Wrapper, h_cvectr_mult.c:841-1005;
// h3d_diag, mb_ham_spds_micro.c:1453-1505;
// h3d_diag_dangling_bond_shift_zincblende, mb_ham_spds_micro.c:1910-2152).
#include "nh_internal.h"
#include "nh_sk.h"

namespace nh {
namespace {

// Strained bond seen from the cation: nnv, nnv0 point anion -> cation
// (getRelativePosVec, geo_struct.c:2278-2294, with the cation-row sign
// flip of h_cvectr_mult.c:1073-1090).
__device__ __forceinline__ void cation_bond_vectors(
    const int32_t *bond_q, const double *disp, const double *lat, int i,
    int slot, int j, double a0, double nnv[3], double nnv0[3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double nrm = 0.25 * bond_q[(i * 4 + slot) * 3 + k];
    nnv0[k] = -(a0 * nrm);
    nnv[k] = -(lat[k] * nrm + disp[j * 3 + k] - disp[i * 3 + k]);
  }
}

__global__ void k_bond(int n_cations, int S, const int32_t *__restrict__ cations,
                       const int32_t *__restrict__ species,
                       const int32_t *__restrict__ nbr,
                       const int32_t *__restrict__ rev_slot,
                       const int32_t *__restrict__ bond_q,
                       const double *__restrict__ disp, double lx, double ly,
                       double lz, const double *__restrict__ parmat,
                       const double *__restrict__ strnmat,
                       double *__restrict__ blocks, double *__restrict__ corr,
                       double *__restrict__ sk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_cations * 4) return;
  const int ci = t >> 2, slot = t & 3;
  const int i = cations[ci];
  const int j = nbr[i * 4 + slot];
  if (j < 0) return;
  const int jslot = rev_slot[i * 4 + slot];
  const double *par = parmat + ((size_t)species[i] * S + species[j]) * 35;
  const double *strn = strnmat + ((size_t)species[i] * S + species[j]) * 33;
  const double lat[3] = {lx, ly, lz};
  double nnv[3], nnv0[3];
  cation_bond_vectors(bond_q, disp, lat, i, slot, j, par[P_A0], nnv, nnv0);
  const BondGeom g = bond_geom(nnv, nnv0, par[P_A0]);

  double v[21];
  scale_integrals(par, strn, g.brat, v);
  double T[100];
  sk_table(g.l, g.m, g.n, v, T);
  if (blocks) {  // stored-block layout only
    double *Ba = blocks + ((size_t)j * 4 + jslot) * 100;  // anion row: T
    double *Bc = blocks + ((size_t)i * 4 + slot) * 100;   // cation row: T^T
#pragma unroll
    for (int a = 0; a < 10; ++a)
#pragma unroll
      for (int c = 0; c < 10; ++c) {
        Ba[a * 10 + c] = T[a * 10 + c];
        Bc[c * 10 + a] = T[a * 10 + c];
      }
  }
  if (sk) {
    double *rec = sk + (size_t)t * kSkStride;
    rec[0] = g.l;
    rec[1] = g.m;
    rec[2] = g.n;
#pragma unroll
    for (int k = 0; k < 21; ++k) rec[3 + k] = v[k];
  }
  // Boykin terms ((V0^2 - V^2) c_strain / (E_a + E_c)) * c_pair
  // (mb_ham_spds_VoffdEdiag_constructors.h; the pVpE_* macros are not
  // parenthesised, so the pair coefficient multiplies the quotient).
  double T0[100];
  sk_table(g.l0, g.m0, g.n0, par + 10, T0);
  const double cs = strn[S_CONST];
  double cc[10];
#pragma unroll
  for (int c = 0; c < 10; ++c) cc[c] = 0.0;
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    double ca = 0.0;
#pragma unroll
    for (int c = 0; c < 10; ++c) {
      const int sa = shell_of(a), sc = shell_of(c);
      const double r1 = T[a * 10 + c], r2 = T0[a * 10 + c];
      const double r3 = (r2 * r2 - r1 * r1) * cs /
                        (par[e_anion(sa)] + par[e_cation(sc)]) *
                        strn[cpair_index(sa, sc)];
      ca += r3;
      cc[c] += r3;
    }
    corr[((size_t)j * 4 + jslot) * 10 + a] = ca;
  }
#pragma unroll
  for (int c = 0; c < 10; ++c) corr[((size_t)i * 4 + slot) * 10 + c] = cc[c];
}

__global__ void k_onsite(int A, int S, const int32_t *__restrict__ species,
                         const int32_t *__restrict__ is_cation,
                         const int32_t *__restrict__ nbr,
                         const int32_t *__restrict__ bond_q,
                         const double *__restrict__ parmat,
                         const double *__restrict__ strnmat,
                         const double *__restrict__ phi, int passivate,
                         double db_shift, const double *__restrict__ corr,
                         double *__restrict__ onsite) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A) return;
  const int ti = species[i];
  const bool cat = is_cation[i] != 0;
  // averaged parameters over the present neighbours (only the onsite
  // energies and the spin-orbit constants are used)
  double e[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) e[k] = 0.0;
  double strain_shift = 0.0;
  int cnt = 0;
  for (int n = 0; n < 4; ++n) {
    const int j = nbr[i * 4 + n];
    if (j < 0) continue;
    const int tj = species[j];
    const int c = cat ? ti : tj, a = cat ? tj : ti;
    const double *p = parmat + ((size_t)c * S + a) * 35;
#pragma unroll
    for (int k = 0; k < 10; ++k) e[k] += p[k];
    strain_shift +=
        strnmat[((size_t)c * S + a) * 33 + (cat ? S_CAT_SHIFT : S_AN_SHIFT)];
    ++cnt;
  }
  strain_shift /= cnt;
#pragma unroll
  for (int k = 0; k < 10; ++k) e[k] /= cnt;
  double hd[10];
#pragma unroll
  for (int o = 0; o < 10; ++o)
    hd[o] = e[cat ? e_cation(shell_of(o)) : e_anion(shell_of(o))];
  const double lambda = cat ? e[P_LC] : e[P_LA];
  double sp[6] = {0, 0, 0, 0, 0, 0};
  if (passivate && cnt < 4) {
    // sp3 hybrid directions of the dangling-bond projector
    const double hyb[4][3] = {{0.25, 0.25, 0.25},
                              {-0.25, -0.25, 0.25},
                              {0.25, -0.25, -0.25},
                              {-0.25, 0.25, -0.25}};
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = 0.25 * db_shift;
    double cosv = 0.0;
    for (int n = 0; n < 4; ++n) {
      double d[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) d[k] = 0.25 * bond_q[(i * 4 + n) * 3 + k];
      int ib = 0;
      cosv = d[0] * hyb[0][0] + d[1] * hyb[0][1] + d[2] * hyb[0][2];
      while (fabs(cosv) < 0.15 && ib < 3) {
        ++ib;
        cosv = d[0] * hyb[ib][0] + d[1] * hyb[ib][1] + d[2] * hyb[ib][2];
      }
      if (nbr[i * 4 + n] >= 0) v[ib] = 0.0;
    }
    const double s = cosv > 0 ? 1.0 : -1.0;
    const double dsp = v[0] + v[1] + v[2] + v[3];
    sp[0] = s * (v[0] - v[1] + v[2] - v[3]);
    sp[1] = s * (v[0] - v[1] - v[2] + v[3]);
    sp[2] = s * (v[0] + v[1] - v[2] - v[3]);
    sp[3] = v[0] + v[1] - v[2] - v[3];
    sp[4] = v[0] - v[1] - v[2] + v[3];
    sp[5] = v[0] - v[1] + v[2] - v[3];
#pragma unroll
    for (int k = 1; k < 5; ++k) hd[k] += dsp;
  }
  const double ph = phi ? phi[i] : 0.0;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    hd[k] += strain_shift;
    hd[k] -= ph;
  }
  for (int n = 0; n < 4; ++n) {
    if (nbr[i * 4 + n] < 0) continue;
#pragma unroll
    for (int k = 0; k < 10; ++k) hd[k] += corr[((size_t)i * 4 + n) * 10 + k];
  }
  double *os = onsite + (size_t)i * kOnsiteStride;
#pragma unroll
  for (int k = 0; k < 10; ++k) os[k] = hd[k];
#pragma unroll
  for (int k = 0; k < 6; ++k) os[10 + k] = sp[k];
  os[16] = lambda;
  os[17] = 0.0;
}

// Matrix-free layout tables: per atom its displacement and species, per
// slot the ideal bond vector seen from the bond's cation (cation rows keep
// their own vector, anion rows negate theirs), packed as 3 x int8.
__global__ void k_mf_geo(int A, const int32_t *__restrict__ species,
                         const int32_t *__restrict__ is_cation,
                         const int32_t *__restrict__ bond_q,
                         const double *__restrict__ disp,
                         double4 *__restrict__ geo, int32_t *__restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A) return;
  geo[i] = make_double4(disp[i * 3], disp[i * 3 + 1], disp[i * 3 + 2],
                        (double)species[i]);
  const int sg = is_cation[i] ? 1 : -1;
  for (int n = 0; n < 4; ++n) {
    int packed = 0;
    for (int k = 0; k < 3; ++k)
      packed |= ((sg * bond_q[(i * 4 + n) * 3 + k]) & 0xff) << (8 * k);
    q[i * 4 + n] = packed;
  }
}

// Per (cation, anion) species pair: integrals 10..30, strain exponents
// 0..14, lattice constant (the parameters k_bond reads).
__global__ void k_mf_pair(int S, const double *__restrict__ parmat,
                          const double *__restrict__ strnmat,
                          double *__restrict__ pair) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= S * S) return;
  const double *par = parmat + (size_t)p * 35;
  const double *strn = strnmat + (size_t)p * 33;
  double *o = pair + (size_t)p * kMfPair;
  for (int k = 0; k < 21; ++k) o[k] = par[10 + k];
  for (int e = 0; e < 15; ++e) o[kMfEta + e] = strn[e];
  o[kMfA0] = par[P_A0];
  for (int k = kMfA0 + 1; k < kMfPair; ++k) o[k] = 0.0;
}

template <typename T>
__global__ void k_convert(int64_t n, const double *__restrict__ src,
                          T *__restrict__ dst) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    dst[k] = T(src[k]);
}

inline int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

static void free_ham(nh_context *c) {
  (void)hipFree(c->onsite);
  (void)hipFree(c->blocks64);
  (void)hipFree(c->blocks32);
  (void)hipFree(c->onsite32);
  (void)hipFree(c->sk64);
  (void)hipFree(c->sk32);
  (void)hipFree(c->corr);
  (void)hipFree(c->mf_geo);
  (void)hipFree(c->mf_q);
  (void)hipFree(c->mf_pair);
  c->mf_geo = nullptr;
  c->mf_q = nullptr;
  c->mf_pair = nullptr;
  c->onsite = nullptr;
  c->blocks64 = nullptr;
  c->blocks32 = nullptr;
  c->onsite32 = nullptr;
  c->sk64 = nullptr;
  c->sk32 = nullptr;
  c->corr = nullptr;
  c->layout = c->precision = -1;
}

nh_status assemble(nh_context *c, int layout, int precision) {
  NH_CHECK(layout == NH_LAYOUT_BLOCKS || layout == NH_LAYOUT_SK ||
               layout == NH_LAYOUT_MATRIX_FREE,
           NH_ERR_ARG, "layout %d not available", layout);
  NH_CHECK(precision == NH_PREC_FP64 || precision == NH_PREC_FP32, NH_ERR_ARG,
           "precision %d not available", precision);
  NH_HIP(hipStreamSynchronize(c->stream));
  free_ham(c);
  const int A = c->n_atoms;
  const size_t nblk = (size_t)A * 4 * 100;
  NH_TRY(device_malloc(c, (void **)&c->onsite, (size_t)A * kOnsiteStride * sizeof(double)));
  NH_TRY(device_malloc(c, (void **)&c->corr, (size_t)A * 4 * 10 * sizeof(double)));
  NH_HIP(hipMemsetAsync(c->corr, 0, (size_t)A * 40 * sizeof(double), c->stream));
  if (layout == NH_LAYOUT_BLOCKS) {
    NH_TRY(device_malloc(c, (void **)&c->blocks64, nblk * sizeof(double)));
    NH_HIP(hipMemsetAsync(c->blocks64, 0, nblk * sizeof(double), c->stream));
  }
  if (layout == NH_LAYOUT_SK) {
    NH_TRY(device_malloc(c, (void **)&c->sk64, (size_t)c->n_cations * 4 * kSkStride * sizeof(double)));
    NH_HIP(hipMemsetAsync(c->sk64, 0,
                          (size_t)c->n_cations * 4 * kSkStride * sizeof(double),
                          c->stream));
  }
  const int S = c->n_species;
  if (layout == NH_LAYOUT_MATRIX_FREE) {
    NH_TRY(device_malloc(c, (void **)&c->mf_geo, (size_t)A * sizeof(double4)));
    NH_TRY(device_malloc(c, (void **)&c->mf_q, (size_t)A * 4 * sizeof(int32_t)));
    NH_TRY(device_malloc(c, (void **)&c->mf_pair, (size_t)S * S * kMfPair * sizeof(double)));
    hipLaunchKernelGGL(k_mf_geo, dim3(grid_for(A, 128)), dim3(128), 0, c->stream,
                       A, c->species, c->is_cation, c->bond_q, c->disp,
                       c->mf_geo, c->mf_q);
    hipLaunchKernelGGL(k_mf_pair, dim3(1), dim3(64 * ((S * S + 63) / 64)), 0,
                       c->stream, S, c->parmat, c->strnmat, c->mf_pair);
    NH_HIP(hipGetLastError());
  }
  // bonds: the stored blocks or SK records (if kept) and the Boykin terms
  // of the onsite diagonal, which every layout stores
  hipLaunchKernelGGL(k_bond, dim3(grid_for((int64_t)c->n_cations * 4, 128)),
                     dim3(128), 0, c->stream, c->n_cations, S, c->cations,
                     c->species, c->nbr, c->rev_slot, c->bond_q, c->disp,
                     c->lattice[0], c->lattice[1], c->lattice[2], c->parmat,
                     c->strnmat, c->blocks64, c->corr, c->sk64);
  NH_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_onsite, dim3(grid_for(A, 128)), dim3(128), 0, c->stream,
                     A, S, c->species, c->is_cation, c->nbr, c->bond_q,
                     c->parmat, c->strnmat, c->phi, c->passivate, c->db_shift,
                     c->corr, c->onsite);
  NH_HIP(hipGetLastError());
  NH_HIP(hipStreamSynchronize(c->stream));
  (void)hipFree(c->corr);  // only the assembly needs the per-bond terms
  c->corr = nullptr;
  c->layout = layout;
  c->precision = NH_PREC_FP64;
  return select_precision(c, precision);
}

nh_status select_precision(nh_context *c, int precision) {
  NH_CHECK(c->layout == NH_LAYOUT_BLOCKS || sk_like(c), NH_ERR_STATE,
           "build the Hamiltonian first");
  NH_CHECK(precision == NH_PREC_FP64 || precision == NH_PREC_FP32, NH_ERR_ARG,
           "precision %d not available", precision);
  const int A = c->n_atoms;
  if (precision == NH_PREC_FP32) {
    // the FP64 values are always kept: derive the FP32 copies once (the
    // matrix-free layout needs none)
    if (c->layout == NH_LAYOUT_BLOCKS && !c->onsite32) {
      NH_TRY(device_malloc(c, (void **)&c->onsite32, (size_t)A * kOnsiteStride * sizeof(float)));
      hipLaunchKernelGGL(k_convert<float>, dim3(1024), dim3(256), 0, c->stream,
                         (int64_t)A * kOnsiteStride, c->onsite, c->onsite32);
    }
    if (c->layout == NH_LAYOUT_BLOCKS && !c->blocks32) {
      const size_t nblk = (size_t)A * 4 * 100;
      NH_TRY(device_malloc(c, (void **)&c->blocks32, nblk * sizeof(float)));
      hipLaunchKernelGGL(k_convert<float>, dim3(4096), dim3(256), 0, c->stream,
                         (int64_t)nblk, c->blocks64, c->blocks32);
    }
    if (c->layout == NH_LAYOUT_SK && !c->sk32) {
      const int64_t ns = (int64_t)c->n_cations * 4 * kSkStride;
      NH_TRY(device_malloc(c, (void **)&c->sk32, ns * sizeof(float)));
      hipLaunchKernelGGL(k_convert<float>, dim3(1024), dim3(256), 0, c->stream,
                         ns, c->sk64, c->sk32);
    }
    NH_HIP(hipGetLastError());
    NH_HIP(hipStreamSynchronize(c->stream));
  }
  c->precision = precision;
  return NH_OK;
}

}  // namespace nh
