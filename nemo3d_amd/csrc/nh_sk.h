// Device-side sp3d5s* Slater-Koster table, strain scaling and Boykin
// diagonal corrections for one zincblende bond.
//
// Orientation: T[a*10 + c] = <a on anion | H | c on cation> along the
// direction cosines (l, m, n) pointing anion -> cation.  Orbital order
// s*, s, px, py, pz, dxy, dyz, dzx, dx2-y2, dz2 (mb_ham_spds_micro.h:97-116).
// Parameter indices follow mb_ham_spds_micro.h:118-264.
#pragma once

#include <hip/hip_runtime.h>

namespace nh {

// sqrt(3) as written by the reference (mb_ham_spds_micro.c:42); it
// differs from the true value by 1.2e-12 relative and enters every d entry
// and the ideal bond length, so the kernels keep it for parity.
constexpr double kSqrt3Ref = 1.73205080757;

enum : int {
  P_E_SA = 0, P_E_PA = 1, P_E_SC = 2, P_E_PC = 3, P_E_STA = 4, P_E_STC = 5,
  P_E_DA = 6, P_E_DC = 7, P_LA = 8, P_LC = 9, P_SSS = 10, P_STSTS = 11,
  P_STASS = 12, P_SASTS = 13, P_SAPCS = 14, P_SCPAS = 15, P_STAPCS = 16,
  P_STCPAS = 17, P_SADCS = 18, P_SCDAS = 19, P_STADCS = 20, P_STCDAS = 21,
  P_PPS = 22, P_PPP = 23, P_PADCS = 24, P_PCDAS = 25, P_PADCP = 26,
  P_PCDAP = 27, P_DDS = 28, P_DDP = 29, P_DDD = 30, P_A0 = 31
};
enum : int { S_CONST = 10, S_CAT_SHIFT = 15, S_AN_SHIFT = 16 };

// The 21 two-centre integrals in parameter order 10..30 and the strain
// exponent index of each (strain_corr_prep, mb_ham_spds_micro.c:2848-2978).
__host__ __device__ constexpr int eta_index(int v) {
  // v = 0..20 <-> parameter 10..30
  return v == 0 ? 2 : v == 1 ? 1 : (v == 2 || v == 3) ? 0
       : (v == 4 || v == 5) ? 3 : (v == 6 || v == 7) ? 7
       : (v == 8 || v == 9) ? 6 : (v == 10 || v == 11) ? 14
       : v == 12 ? 4 : v == 13 ? 5 : (v == 14 || v == 15) ? 8
       : (v == 16 || v == 17) ? 9 : v == 18 ? 11 : v == 19 ? 12 : 13;
}

// Shell of each orbital (0 s*, 1 s, 2 p, 3 d).
__host__ __device__ constexpr int shell_of(int o) {
  return o == 0 ? 0 : o == 1 ? 1 : o < 5 ? 2 : 3;
}
// Boykin coefficient index for (anion shell, cation shell)
// (pVpE_* pairing, mb_ham_spds_micro.c:100-143).
__host__ __device__ constexpr int cpair_index(int sa, int sc) {
  constexpr int t[16] = {18, 19, 23, 27, 20, 17, 21, 25,
                         24, 22, 29, 30, 28, 26, 31, 32};
  return t[sa * 4 + sc];
}
__host__ __device__ constexpr int e_anion(int s) {
  return s == 0 ? P_E_STA : s == 1 ? P_E_SA : s == 2 ? P_E_PA : P_E_DA;
}
__host__ __device__ constexpr int e_cation(int s) {
  return s == 0 ? P_E_STC : s == 1 ? P_E_SC : s == 2 ? P_E_PC : P_E_DC;
}

// Slater-Koster table (Slater & Koster 1954, Table I; sp3d5s* form of
// Jancu et al. 1998).  v points at the 21 integrals (parameters 10..30);
// the integrals with the lower angular momentum on the cation enter with a
// minus sign, as in the reference (mb_ham_spds_micro.c:58-78).
template <typename R>
__device__ __forceinline__ void sk_table(R l, R m, R n, const R *v, R *T) {
  const R s3 = R(kSqrt3Ref);
  const R l2 = l * l, m2 = m * m, n2 = n * n;
  const R sss = v[0], ststs = v[1], stass = v[2], sasts = v[3], sapcs = v[4],
          scpas = -v[5], stapcs = v[6], stcpas = -v[7], sadcs = v[8],
          scdas = v[9], stadcs = v[10], stcdas = v[11], pps = v[12],
          ppp = v[13], padcs = v[14], pcdas = -v[15], padcp = v[16],
          pcdap = -v[17], dds = v[18], ddp = v[19], ddd = v[20];
  const R q[3] = {l, m, n};
  const R sd[5] = {s3 * l * m, s3 * m * n, s3 * n * l, R(0.5) * s3 * (l2 - m2),
                   n2 - R(0.5) * (l2 + m2)};
  T[0] = ststs;
  T[1] = stass;
  T[10] = sasts;
  T[11] = sss;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T[2 + c] = q[c] * stapcs;
    T[12 + c] = q[c] * sapcs;
    T[(2 + c) * 10 + 0] = q[c] * stcpas;
    T[(2 + c) * 10 + 1] = q[c] * scpas;
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    T[5 + c] = sd[c] * stadcs;
    T[15 + c] = sd[c] * sadcs;
    T[(5 + c) * 10 + 0] = sd[c] * stcdas;
    T[(5 + c) * 10 + 1] = sd[c] * scdas;
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c)
      T[(2 + a) * 10 + 2 + c] =
          a == c ? q[a] * q[a] * pps + (R(1) - q[a] * q[a]) * ppp
                 : q[a] * q[c] * pps - q[a] * q[c] * ppp;
  const R lmn = l * m * n;
  const R pd_s[3][5] = {
      {s3 * l2 * m, s3 * lmn, s3 * l2 * n, R(0.5) * s3 * l * (l2 - m2),
       l * (n2 - R(0.5) * (l2 + m2))},
      {s3 * m2 * l, s3 * m2 * n, s3 * lmn, R(0.5) * s3 * m * (l2 - m2),
       m * (n2 - R(0.5) * (l2 + m2))},
      {s3 * lmn, s3 * n2 * m, s3 * n2 * l, R(0.5) * s3 * n * (l2 - m2),
       n * (n2 - R(0.5) * (l2 + m2))}};
  const R pd_p[3][5] = {
      {m * (R(1) - R(2) * l2), R(-2) * lmn, n * (R(1) - R(2) * l2),
       l * (R(1) - l2 + m2), -s3 * l * n2},
      {l * (R(1) - R(2) * m2), n * (R(1) - R(2) * m2), R(-2) * lmn,
       -m * (R(1) + l2 - m2), -s3 * m * n2},
      {R(-2) * lmn, m * (R(1) - R(2) * n2), l * (R(1) - R(2) * n2),
       -n * (l2 - m2), s3 * n * (l2 + m2)}};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      T[(2 + a) * 10 + 5 + c] = pd_s[a][c] * padcs + pd_p[a][c] * padcp;
      T[(5 + c) * 10 + 2 + a] = pd_s[a][c] * pcdas + pd_p[a][c] * pcdap;
    }
  const R lm = l * m, mn = m * n, nl = n * l;
  const R dl = l2 - m2, z = n2 - R(0.5) * (l2 + m2), lpm = l2 + m2;
  R dd[15];
  dd[0] = R(3) * l2 * m2 * dds + (l2 + m2 - R(4) * l2 * m2) * ddp +
          (n2 + l2 * m2) * ddd;
  dd[1] = R(3) * l * m2 * n * dds + nl * (R(1) - R(4) * m2) * ddp +
          nl * (m2 - R(1)) * ddd;
  dd[2] = R(3) * l2 * mn * dds + mn * (R(1) - R(4) * l2) * ddp +
          mn * (l2 - R(1)) * ddd;
  dd[3] = R(1.5) * lm * dl * dds + R(2) * lm * (m2 - l2) * ddp +
          R(0.5) * lm * dl * ddd;
  dd[4] = s3 * lm * z * dds - s3 * R(2) * lm * n2 * ddp +
          R(0.5) * s3 * lm * (R(1) + n2) * ddd;
  dd[5] = R(3) * m2 * n2 * dds + (m2 + n2 - R(4) * m2 * n2) * ddp +
          (l2 + m2 * n2) * ddd;
  dd[6] = R(3) * m * n2 * l * dds + l * m * (R(1) - R(4) * n2) * ddp +
          l * m * (n2 - R(1)) * ddd;
  dd[7] = R(1.5) * mn * dl * dds - mn * (R(1) + R(2) * dl) * ddp +
          mn * (R(1) + R(0.5) * dl) * ddd;
  dd[8] = s3 * mn * z * dds + s3 * mn * (lpm - n2) * ddp -
          R(0.5) * s3 * mn * lpm * ddd;
  dd[9] = R(3) * n2 * l2 * dds + (n2 + l2 - R(4) * n2 * l2) * ddp +
          (m2 + n2 * l2) * ddd;
  dd[10] = R(1.5) * nl * dl * dds + nl * (R(1) - R(2) * dl) * ddp -
           nl * (R(1) - R(0.5) * dl) * ddd;
  dd[11] = s3 * nl * z * dds + s3 * nl * (lpm - n2) * ddp -
           R(0.5) * s3 * nl * lpm * ddd;
  dd[12] = R(0.75) * dl * dl * dds + (lpm - dl * dl) * ddp +
           (n2 + R(0.25) * dl * dl) * ddd;
  dd[13] = R(0.5) * s3 * dl * z * dds + s3 * n2 * (m2 - l2) * ddp +
           R(0.25) * s3 * (R(1) + n2) * dl * ddd;
  dd[14] = z * z * dds + R(3) * n2 * lpm * ddp + R(0.75) * lpm * lpm * ddd;
  // packed upper triangle index of (a, c), a <= c, in the 5x5 d block
  constexpr int up[5][5] = {{0, 1, 2, 3, 4},
                            {1, 5, 6, 7, 8},
                            {2, 6, 9, 10, 11},
                            {3, 7, 10, 12, 13},
                            {4, 8, 11, 13, 14}};
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int c = 0; c < 5; ++c) T[(5 + a) * 10 + 5 + c] = dd[up[a][c]];
}

// Direction cosines and bond-length ratio of one bond
// (strain_corr_prep, mb_ham_spds_micro.c:2855-2879).
struct BondGeom {
  double l, m, n, l0, m0, n0, brat;
};

__device__ __forceinline__ BondGeom bond_geom(const double nnv[3],
                                              const double nnv0[3],
                                              double a0) {
  BondGeom g;
  const double den0 =
      sqrt(nnv0[0] * nnv0[0] + nnv0[1] * nnv0[1] + nnv0[2] * nnv0[2]);
  g.l0 = nnv0[0] / den0;
  g.m0 = nnv0[1] / den0;
  g.n0 = nnv0[2] / den0;
  const double den = sqrt(nnv[0] * nnv[0] + nnv[1] * nnv[1] + nnv[2] * nnv[2]);
  g.l = nnv[0] / den;
  g.m = nnv[1] / den;
  g.n = nnv[2] / den;
  g.brat = (0.25 * kSqrt3Ref * a0) / den;
  return g;
}

// Strain-scaled integrals V_k * brat^eta_k (parameters 10..30).
__device__ __forceinline__ void scale_integrals(const double *par,
                                                const double *strn,
                                                double brat, double *v) {
#pragma unroll
  for (int k = 0; k < 21; ++k) v[k] = par[10 + k] * pow(brat, strn[eta_index(k)]);
}

}  // namespace nh
