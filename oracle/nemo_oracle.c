/*
 * CPU oracle for the NEMO-3D sp3d5s* eigensolver hot path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference
 * algorithm used as the checker by tests/, smoke() and the cpu_baseline
 * leg of bench.py.  The product library never links or calls it.
 * Each function cites the reference file:line it restates (paths relative
 * to NEMO_3D/).  How each function is pinned is listed in nemo_oracle.h.
 */
#include "nemo_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* The reference writes sqrt(3) as this truncated literal
 * (nemo3d/src/ham/mb_ham_spds_micro.c:42).  It enters every d-orbital
 * Slater-Koster entry and the bond-length ratio, and differs from the true
 * value by 1.2e-12 relative, so the oracle keeps it. */
#define OR_SQRT3 1.73205080757

/* Parameter vector indices (nemo3d/src/ham/mb_ham_spds_micro.h:118-219). */
enum {
  P_E_SA = 0, P_E_PA = 1, P_E_SC = 2, P_E_PC = 3, P_E_STA = 4, P_E_STC = 5,
  P_E_DA = 6, P_E_DC = 7, P_LA = 8, P_LC = 9,
  P_SSS = 10, P_STSTS = 11, P_STASS = 12, P_SASTS = 13, P_SAPCS = 14,
  P_SCPAS = 15, P_STAPCS = 16, P_STCPAS = 17, P_SADCS = 18, P_SCDAS = 19,
  P_STADCS = 20, P_STCDAS = 21, P_PPS = 22, P_PPP = 23, P_PADCS = 24,
  P_PCDAS = 25, P_PADCP = 26, P_PCDAP = 27, P_DDS = 28, P_DDP = 29,
  P_DDD = 30, P_A0 = 31
};
/* Strain vector indices (nemo3d/src/ham/mb_ham_spds_micro.h:226-264). */
enum {
  S_ETA_SSTS = 0, S_ETA_STSTS = 1, S_ETA_SSS = 2, S_ETA_SPS = 3,
  S_ETA_PPS = 4, S_ETA_PPP = 5, S_ETA_SDS = 6, S_ETA_STPS = 7,
  S_ETA_PDS = 8, S_ETA_PDP = 9, S_CONST = 10, S_ETA_DDS = 11,
  S_ETA_DDP = 12, S_ETA_DDD = 13, S_ETA_STDS = 14, S_CAT_SHIFT = 15,
  S_AN_SHIFT = 16
};

/* Shell of each of the 10 orbitals: 0 s*, 1 s, 2 p, 3 d
 * (orbital order s*, s, px, py, pz, dxy, dyz, dzx, dx2-y2, dz2:
 * nemo3d/src/ham/mb_ham_spds_micro.h:97-116). */
static const int kShell[10] = {0, 1, 2, 2, 2, 3, 3, 3, 3, 3};

/* Boykin coefficient index for (anion shell, cation shell); the pairing
 * follows the pVpE_* definitions of mb_ham_spds_micro.c:100-143. */
static const int kCPair[4][4] = {
    /* anion s*  */ {18, 19, 23, 27},
    /* anion s   */ {20, 17, 21, 25},
    /* anion p   */ {24, 22, 29, 30},
    /* anion d   */ {28, 26, 31, 32}};

/* Onsite energy index of each shell for the anion and for the cation. */
static const int kEAnion[4] = {P_E_STA, P_E_SA, P_E_PA, P_E_DA};
static const int kECation[4] = {P_E_STC, P_E_SC, P_E_PC, P_E_DC};

/* Slater-Koster table (Slater & Koster, Phys. Rev. 94, 1498 (1954),
 * Table I; sp3d5s* form of Jancu et al., PRB 57, 6493 (1998)) for an
 * anion orbital alpha (row) and a cation orbital beta (column) along the
 * direction cosines (l, m, n) pointing anion -> cation.  The sign
 * convention of the two-centre integrals follows the reference: the
 * integrals with the cation orbital of lower angular momentum enter with a
 * minus sign (scpas, stcpas, pcdas, pcdap in mb_ham_spds_micro.c:58-78). */
static void sk_table(double l, double m, double n, const double *v,
                     double T[100]) {
  const double s3 = OR_SQRT3;
  const double l2 = l * l, m2 = m * m, n2 = n * n;
  const double sss = v[P_SSS], ststs = v[P_STSTS], stass = v[P_STASS],
               sasts = v[P_SASTS], sapcs = v[P_SAPCS], scpas = -v[P_SCPAS],
               stapcs = v[P_STAPCS], stcpas = -v[P_STCPAS],
               sadcs = v[P_SADCS], scdas = v[P_SCDAS], stadcs = v[P_STADCS],
               stcdas = v[P_STCDAS], pps = v[P_PPS], ppp = v[P_PPP],
               padcs = v[P_PADCS], pcdas = -v[P_PCDAS], padcp = v[P_PADCP],
               pcdap = -v[P_PCDAP], dds = v[P_DDS], ddp = v[P_DDP],
               ddd = v[P_DDD];
  const double q[3] = {l, m, n};
  /* s-d angular factors for xy, yz, zx, x2-y2, z2 */
  const double sd[5] = {s3 * l * m, s3 * m * n, s3 * n * l,
                        0.5 * s3 * (l2 - m2), n2 - 0.5 * (l2 + m2)};
  int a, c;

  /* rows s*, s */
  T[0 * 10 + 0] = ststs;
  T[0 * 10 + 1] = stass;
  T[1 * 10 + 0] = sasts;
  T[1 * 10 + 1] = sss;
  for (c = 0; c < 3; ++c) {
    T[0 * 10 + 2 + c] = q[c] * stapcs;
    T[1 * 10 + 2 + c] = q[c] * sapcs;
    /* p rows with s-type cation columns */
    T[(2 + c) * 10 + 0] = q[c] * stcpas;
    T[(2 + c) * 10 + 1] = q[c] * scpas;
  }
  for (c = 0; c < 5; ++c) {
    T[0 * 10 + 5 + c] = sd[c] * stadcs;
    T[1 * 10 + 5 + c] = sd[c] * sadcs;
    T[(5 + c) * 10 + 0] = sd[c] * stcdas;
    T[(5 + c) * 10 + 1] = sd[c] * scdas;
  }
  /* p-p */
  for (a = 0; a < 3; ++a)
    for (c = 0; c < 3; ++c) {
      if (a == c)
        T[(2 + a) * 10 + 2 + c] = q[a] * q[a] * pps + (1 - q[a] * q[a]) * ppp;
      else
        T[(2 + a) * 10 + 2 + c] = q[a] * q[c] * pps - q[a] * q[c] * ppp;
    }
  /* p-d: pd[p][d] with sigma and pi integrals */
  {
    double pd_s[3][5], pd_p[3][5];
    const double lmn = l * m * n;
    pd_s[0][0] = s3 * l2 * m;            pd_p[0][0] = m * (1 - 2 * l2);
    pd_s[0][1] = s3 * lmn;               pd_p[0][1] = -2 * lmn;
    pd_s[0][2] = s3 * l2 * n;            pd_p[0][2] = n * (1 - 2 * l2);
    pd_s[0][3] = 0.5 * s3 * l * (l2 - m2); pd_p[0][3] = l * (1 - l2 + m2);
    pd_s[0][4] = l * (n2 - 0.5 * (l2 + m2)); pd_p[0][4] = -s3 * l * n2;
    pd_s[1][0] = s3 * m2 * l;            pd_p[1][0] = l * (1 - 2 * m2);
    pd_s[1][1] = s3 * m2 * n;            pd_p[1][1] = n * (1 - 2 * m2);
    pd_s[1][2] = s3 * lmn;               pd_p[1][2] = -2 * lmn;
    pd_s[1][3] = 0.5 * s3 * m * (l2 - m2); pd_p[1][3] = -m * (1 + l2 - m2);
    pd_s[1][4] = m * (n2 - 0.5 * (l2 + m2)); pd_p[1][4] = -s3 * m * n2;
    pd_s[2][0] = s3 * lmn;               pd_p[2][0] = -2 * lmn;
    pd_s[2][1] = s3 * n2 * m;            pd_p[2][1] = m * (1 - 2 * n2);
    pd_s[2][2] = s3 * n2 * l;            pd_p[2][2] = l * (1 - 2 * n2);
    pd_s[2][3] = 0.5 * s3 * n * (l2 - m2); pd_p[2][3] = -n * (l2 - m2);
    pd_s[2][4] = n * (n2 - 0.5 * (l2 + m2)); pd_p[2][4] = s3 * n * (l2 + m2);
    for (a = 0; a < 3; ++a)
      for (c = 0; c < 5; ++c) {
        T[(2 + a) * 10 + 5 + c] = pd_s[a][c] * padcs + pd_p[a][c] * padcp;
        T[(5 + c) * 10 + 2 + a] = pd_s[a][c] * pcdas + pd_p[a][c] * pcdap;
      }
  }
  /* d-d (symmetric in the two orbitals) */
  {
    double dd[5][5];
    const double lm = l * m, mn = m * n, nl = n * l;
    const double dl = l2 - m2, z = n2 - 0.5 * (l2 + m2), lpm = l2 + m2;
    dd[0][0] = 3 * l2 * m2 * dds + (l2 + m2 - 4 * l2 * m2) * ddp +
               (n2 + l2 * m2) * ddd;
    dd[0][1] = 3 * l * m2 * n * dds + nl * (1 - 4 * m2) * ddp +
               nl * (m2 - 1) * ddd;
    dd[0][2] = 3 * l2 * mn * dds + mn * (1 - 4 * l2) * ddp +
               mn * (l2 - 1) * ddd;
    dd[0][3] = 1.5 * lm * dl * dds + 2 * lm * (m2 - l2) * ddp +
               0.5 * lm * dl * ddd;
    dd[0][4] = s3 * lm * z * dds - s3 * 2 * lm * n2 * ddp +
               0.5 * s3 * lm * (1 + n2) * ddd;
    dd[1][1] = 3 * m2 * n2 * dds + (m2 + n2 - 4 * m2 * n2) * ddp +
               (l2 + m2 * n2) * ddd;
    dd[1][2] = 3 * m * n2 * l * dds + l * m * (1 - 4 * n2) * ddp +
               l * m * (n2 - 1) * ddd;
    dd[1][3] = 1.5 * mn * dl * dds - mn * (1 + 2 * dl) * ddp +
               mn * (1 + 0.5 * dl) * ddd;
    dd[1][4] = s3 * mn * z * dds + s3 * mn * (lpm - n2) * ddp -
               0.5 * s3 * mn * lpm * ddd;
    dd[2][2] = 3 * n2 * l2 * dds + (n2 + l2 - 4 * n2 * l2) * ddp +
               (m2 + n2 * l2) * ddd;
    dd[2][3] = 1.5 * nl * dl * dds + nl * (1 - 2 * dl) * ddp -
               nl * (1 - 0.5 * dl) * ddd;
    dd[2][4] = s3 * nl * z * dds + s3 * nl * (lpm - n2) * ddp -
               0.5 * s3 * nl * lpm * ddd;
    dd[3][3] = 0.75 * dl * dl * dds + (lpm - dl * dl) * ddp +
               (n2 + 0.25 * dl * dl) * ddd;
    dd[3][4] = 0.5 * s3 * dl * z * dds + s3 * n2 * (m2 - l2) * ddp +
               0.25 * s3 * (1 + n2) * dl * ddd;
    dd[4][4] = z * z * dds + 3 * n2 * lpm * ddp + 0.75 * lpm * lpm * ddd;
    for (a = 0; a < 5; ++a)
      for (c = 0; c < 5; ++c)
        T[(5 + a) * 10 + 5 + c] = a <= c ? dd[a][c] : dd[c][a];
  }
}

/* Strain scaling of the two-centre integrals by (d0/d)^eta
 * (nemo3d/src/ham/mb_ham_spds_micro.c:2848-2978, strain_corr_prep). */
static void scale_integrals(const double *par, const double *strn, double brat,
                            double *out) {
  static const int map[21][2] = {
      {P_SSS, S_ETA_SSS},     {P_STSTS, S_ETA_STSTS}, {P_STASS, S_ETA_SSTS},
      {P_SASTS, S_ETA_SSTS},  {P_SAPCS, S_ETA_SPS},   {P_SCPAS, S_ETA_SPS},
      {P_STAPCS, S_ETA_STPS}, {P_STCPAS, S_ETA_STPS}, {P_SADCS, S_ETA_SDS},
      {P_SCDAS, S_ETA_SDS},   {P_STADCS, S_ETA_STDS}, {P_STCDAS, S_ETA_STDS},
      {P_PPS, S_ETA_PPS},     {P_PPP, S_ETA_PPP},     {P_PADCS, S_ETA_PDS},
      {P_PCDAS, S_ETA_PDS},   {P_PADCP, S_ETA_PDP},   {P_PCDAP, S_ETA_PDP},
      {P_DDS, S_ETA_DDS},     {P_DDP, S_ETA_DDP},     {P_DDD, S_ETA_DDD}};
  int k;
  memcpy(out, par, 35 * sizeof(double));
  for (k = 0; k < 21; ++k)
    out[map[k][0]] = par[map[k][0]] * pow(brat, strn ? strn[map[k][1]] : 0.0);
}

void or_sk_bond(const double *par, const double *strn, double a0,
                const double nnv[3], const double nnv0[3], double T[100],
                double corr_a[10], double corr_c[10]) {
  double den0, den, l0, m0, n0, l, m, n, brat, cstrain;
  double vs[35], T0[100];
  int a, c;

  den0 = sqrt(nnv0[0] * nnv0[0] + nnv0[1] * nnv0[1] + nnv0[2] * nnv0[2]);
  l0 = nnv0[0] / den0;
  m0 = nnv0[1] / den0;
  n0 = nnv0[2] / den0;
  den = sqrt(nnv[0] * nnv[0] + nnv[1] * nnv[1] + nnv[2] * nnv[2]);
  l = nnv[0] / den;
  m = nnv[1] / den;
  n = nnv[2] / den;
  brat = (0.25 * OR_SQRT3 * a0) / den;
  scale_integrals(par, strn, brat, vs);
  sk_table(l, m, n, vs, T);
  sk_table(l0, m0, n0, par, T0);
  cstrain = strn ? strn[S_CONST] : 0.0;

  for (a = 0; a < 10; ++a) corr_a[a] = 0.0;
  for (c = 0; c < 10; ++c) corr_c[c] = 0.0;
  if (!strn) return;
  /* Boykin diagonal correction ((V0^2 - V^2) c_strain / (E_a + E_c)) * c_pair
   * for every entry (mb_ham_spds_VoffdEdiag_constructors.h:46-1251; the
   * pVpE_* macros of mb_ham_spds_micro.c:100-143 are not parenthesised, so
   * the pair coefficient multiplies the quotient). */
  for (a = 0; a < 10; ++a)
    for (c = 0; c < 10; ++c) {
      const int sa = kShell[a], sc = kShell[c];
      const double r1 = T[a * 10 + c], r2 = T0[a * 10 + c];
      const double r3 = (r2 * r2 - r1 * r1) * cstrain /
                        (par[kEAnion[sa]] + par[kECation[sc]]) *
                        strn[kCPair[sa][sc]];
      corr_a[a] += r3;
      corr_c[c] += r3;
    }
}

void or_onsite(const double *param, int anion, double diag[10],
               double *lambda) {
  /* h3d_diag, nemo3d/src/ham/mb_ham_spds_micro.c:1453-1505 */
  int o;
  for (o = 0; o < 10; ++o)
    diag[o] = param[anion ? kEAnion[kShell[o]] : kECation[kShell[o]]];
  *lambda = anion ? param[P_LA] : param[P_LC];
}

/* Dangling-bond sp3-hybrid shift of a surface atom
 * (h3d_diag_dangling_bond_shift_zincblende,
 * nemo3d/src/ham/mb_ham_spds_micro.c:1910-2152).  sp receives the s-p
 * couplings (1,2),(1,3),(1,4),(2,3),(2,4),(3,4) and dsp the shift added to
 * the s, px, py, pz diagonal. */
static void dangling_bond(const int *bond_q, const int *nbr, double shift,
                          double *dsp, double sp[6]) {
  static const double hyb[4][3] = {{0.25, 0.25, 0.25},
                                   {-0.25, -0.25, 0.25},
                                   {0.25, -0.25, -0.25},
                                   {-0.25, 0.25, -0.25}};
  double v[4], cosv = 0.0, s;
  int n, ib, k;
  for (k = 0; k < 4; ++k) v[k] = 0.25 * shift;
  for (n = 0; n < 4; ++n) {
    double d[3];
    for (k = 0; k < 3; ++k) d[k] = 0.25 * bond_q[n * 3 + k];
    ib = 0;
    cosv = d[0] * hyb[0][0] + d[1] * hyb[0][1] + d[2] * hyb[0][2];
    while (fabs(cosv) < 0.15 && ib < 3) {
      ++ib;
      cosv = d[0] * hyb[ib][0] + d[1] * hyb[ib][1] + d[2] * hyb[ib][2];
    }
    if (nbr[n] >= 0) v[ib] = 0.0;
  }
  /* the sign pattern is chosen by the cosine of the last neighbour */
  s = cosv > 0 ? 1.0 : -1.0;
  *dsp = v[0] + v[1] + v[2] + v[3];
  sp[0] = s * (v[0] - v[1] + v[2] - v[3]); /* (s,px) */
  sp[1] = s * (v[0] - v[1] - v[2] + v[3]); /* (s,py) */
  sp[2] = s * (v[0] + v[1] - v[2] - v[3]); /* (s,pz) */
  sp[3] = v[0] + v[1] - v[2] - v[3];       /* (px,py) */
  sp[4] = v[0] - v[1] - v[2] + v[3];       /* (px,pz) */
  sp[5] = v[0] - v[1] + v[2] - v[3];       /* (py,pz) */
}

int or_assemble(const or_model *md, double *onsite, double *blocks) {
  const int S = md->n_species;
  int i, n, k, nsurf = 0;
  for (i = 0; i < md->n_atoms; ++i) {
    const int ti = md->species[i];
    const int cat_i = md->is_cation[i];
    double param[35], hd[10], lambda, strain_shift = 0.0;
    double *os = onsite + (size_t)i * OR_ONSITE_STRIDE;
    int nbrcnt = 0, present = 0;

    /* averaged onsite parameters over the present neighbours
     * (H_Diag_Wrapper, nemo3d/src/base/h_cvectr_mult.c:841-918) */
    for (k = 0; k < 35; ++k) param[k] = 0.0;
    for (n = 0; n < 4; ++n) {
      const int j = md->nbr[i * 4 + n];
      int c, a;
      if (j < 0) continue;
      c = cat_i ? ti : md->species[j];
      a = cat_i ? md->species[j] : ti;
      if (md->strnmat)
        strain_shift +=
            md->strnmat[((size_t)c * S + a) * 33 +
                        (cat_i ? S_CAT_SHIFT : S_AN_SHIFT)];
      for (k = 0; k < 35; ++k)
        param[k] += md->parmat[((size_t)c * S + a) * 35 + k];
      ++nbrcnt;
    }
    strain_shift /= nbrcnt;
    for (k = 0; k < 35; ++k) param[k] /= nbrcnt;
    or_onsite(param, !cat_i, hd, &lambda);
    for (k = 0; k < 6; ++k) os[10 + k] = 0.0;
    present = nbrcnt;
    if (md->passivate && present < 4) {
      double dsp;
      dangling_bond(md->bond_q + (size_t)i * 12, md->nbr + (size_t)i * 4,
                    md->db_shift, &dsp, os + 10);
      for (k = 1; k < 5; ++k) hd[k] += dsp;
      ++nsurf;
    }
    for (k = 0; k < 10; ++k) hd[k] += strain_shift;
    /* potential file / fields: hd -= phi (h_cvectr_mult.c:1226-1231) */
    if (md->phi)
      for (k = 0; k < 10; ++k) hd[k] -= md->phi[i];

    /* bonds (H_DiagStrainCorr_OffDiag_Wrapper, h_cvectr_mult.c:1007-1126,
     * getRelativePosVec, nemo3d/src/base/geo_struct.c:2278-2294) */
    for (n = 0; n < 4; ++n) {
      const int j = md->nbr[i * 4 + n];
      double *B = blocks + ((size_t)i * 4 + n) * 100;
      double nnv[3], nnv0[3], T[100], ca[10], cc[10];
      const double *par, *strn;
      double a0;
      int c, a;
      if (j < 0) {
        memset(B, 0, 100 * sizeof(double));
        continue;
      }
      c = cat_i ? ti : md->species[j];
      a = cat_i ? md->species[j] : ti;
      par = md->parmat + ((size_t)c * S + a) * 35;
      strn = md->strnmat ? md->strnmat + ((size_t)c * S + a) * 33 : NULL;
      /* a0 = parmat[this][nbr][pV_unstr_latt] (h_cvectr_mult.c:1022) */
      a0 = md->parmat[((size_t)ti * S + md->species[j]) * 35 + P_A0];
      for (k = 0; k < 3; ++k) {
        const double nrm = 0.25 * md->bond_q[((size_t)i * 4 + n) * 3 + k];
        nnv0[k] = a0 * nrm;
        nnv[k] = md->lattice[k] * nrm + md->disp[(size_t)j * 3 + k] -
                 md->disp[(size_t)i * 3 + k];
        if (cat_i) { /* cation rows see the anion -> cation direction */
          nnv0[k] = -nnv0[k];
          nnv[k] = -nnv[k];
        }
      }
      or_sk_bond(par, strn, a0, nnv, nnv0, T, ca, cc);
      if (cat_i) {
        int r, q;
        for (r = 0; r < 10; ++r)
          for (q = 0; q < 10; ++q) B[r * 10 + q] = T[q * 10 + r];
        for (k = 0; k < 10; ++k) hd[k] += cc[k];
      } else {
        memcpy(B, T, 100 * sizeof(double));
        for (k = 0; k < 10; ++k) hd[k] += ca[k];
      }
    }
    for (k = 0; k < 10; ++k) os[k] = hd[k];
    os[16] = lambda;
  }
  return nsurf;
}

/* Onsite 20x20 (or 10x10) block applied to one atom.  Spin-orbit entries
 * (h3d_diag with mb_ham_spds_so.h; stored pairs HamZincBlende.c:170-183):
 * up rows 0-9, down rows 10-19, p orbitals 2,3,4 / 12,13,14. */
static void onsite_apply(int nb, const double *os, const double *x,
                         double *y) {
  int s, o;
  const int nsp = nb / 10;
  for (s = 0; s < nsp; ++s) {
    const double *xs = x + 20 * s;
    double *ys = y + 20 * s;
    for (o = 0; o < 10; ++o) {
      ys[2 * o] += os[o] * xs[2 * o];
      ys[2 * o + 1] += os[o] * xs[2 * o + 1];
    }
    {
      static const int pr[6][2] = {{1, 2}, {1, 3}, {1, 4},
                                   {2, 3}, {2, 4}, {3, 4}};
      int k;
      for (k = 0; k < 6; ++k) {
        const double h = os[10 + k];
        const int I = pr[k][0], J = pr[k][1];
        if (h == 0.0) continue;
        ys[2 * I] += h * xs[2 * J];
        ys[2 * I + 1] += h * xs[2 * J + 1];
        ys[2 * J] += h * xs[2 * I];
        ys[2 * J + 1] += h * xs[2 * I + 1];
      }
    }
  }
  if (nb == 20) {
    const double lam = os[16];
    /* upper-triangle entries (I, J, re, im); y_I += h x_J, y_J += conj(h) x_I */
    const struct { int I, J; double re, im; } so[6] = {
        {2, 3, 0.0, -lam},  {12, 13, 0.0, lam}, {4, 12, -lam, 0.0},
        {4, 13, 0.0, lam},  {2, 14, lam, 0.0},  {3, 14, 0.0, -lam}};
    int k;
    for (k = 0; k < 6; ++k) {
      const int I = so[k].I, J = so[k].J;
      const double hr = so[k].re, hi = so[k].im;
      y[2 * I] += hr * x[2 * J] - hi * x[2 * J + 1];
      y[2 * I + 1] += hr * x[2 * J + 1] + hi * x[2 * J];
      y[2 * J] += hr * x[2 * I] + hi * x[2 * I + 1];
      y[2 * J + 1] += hr * x[2 * I + 1] - hi * x[2 * I];
    }
  }
}

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
  (void)n;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_hv(int nb, int n_atoms, const double *onsite, const double *blocks,
           const int *nbr, const double *phase, double shift, const double *x,
           double *y) {
  int i;
  const int nsp = nb / 10;
  /* every atom's rows are computed independently (row-owner layout), so
   * the result does not depend on the thread count */
#pragma omp parallel for schedule(static)
  for (i = 0; i < n_atoms; ++i) {
    int n, s, o, k;
    const double *xi = x + (size_t)i * nb * 2;
    double *yi = y + (size_t)i * nb * 2;
    for (k = 0; k < 2 * nb; ++k) yi[k] = 0.0;
    onsite_apply(nb, onsite + (size_t)i * OR_ONSITE_STRIDE, xi, yi);
    for (n = 0; n < 4; ++n) {
      const int j = nbr[i * 4 + n];
      const double *B = blocks + ((size_t)i * 4 + n) * 100;
      const double *xj;
      double pr = 1.0, pi = 0.0;
      if (j < 0) continue;
      xj = x + (size_t)j * nb * 2;
      if (phase) {
        pr = phase[((size_t)i * 4 + n) * 2];
        pi = phase[((size_t)i * 4 + n) * 2 + 1];
      }
      for (s = 0; s < nsp; ++s)
        for (o = 0; o < 10; ++o) {
          double ar = 0.0, ai = 0.0;
          for (k = 0; k < 10; ++k) {
            ar += B[o * 10 + k] * xj[2 * (10 * s + k)];
            ai += B[o * 10 + k] * xj[2 * (10 * s + k) + 1];
          }
          yi[2 * (10 * s + o)] += pr * ar - pi * ai;
          yi[2 * (10 * s + o) + 1] += pr * ai + pi * ar;
        }
    }
    for (k = 0; k < nb; ++k) {
      yi[2 * k] -= shift * xi[2 * k];
      yi[2 * k + 1] -= shift * xi[2 * k + 1];
    }
  }
}

void or_lanczos_tridiag(int nb, int n_atoms, const double *onsite,
                        const double *blocks, const int *nbr, const double *r0,
                        int iters, double *alpha, double *beta) {
  /* single_lanczos_iteration, NEMO_3D/eigen/src/lanczos.c:421-757 */
  const size_t N = (size_t)n_atoms * nb;
  double *q = (double *)malloc(2 * N * sizeof(double));
  double *q1 = (double *)calloc(2 * N, sizeof(double));
  double *r = (double *)malloc(2 * N * sizeof(double));
  double b = 1.0;
  size_t j;
  int it;
  /* the sums run over a fixed number of contiguous chunks, added in chunk
   * order: the result does not depend on the thread count */
  enum { NCH = 256 };
  double part[NCH];
  int ch;
  memcpy(r, r0, 2 * N * sizeof(double));
  for (it = 0; it < iters; ++it) {
    const double inv = 1.0 / b;
    double al = 0.0, bb = 0.0, bprev = it ? b : 0.0;
#pragma omp parallel for schedule(static)
    for (j = 0; j < 2 * N; ++j) q[j] = r[j] * inv;
    or_hv(nb, n_atoms, onsite, blocks, nbr, NULL, 0.0, q, r);
#pragma omp parallel for schedule(static)
    for (ch = 0; ch < NCH; ++ch) {
      size_t k, k0 = N * ch / NCH, k1 = N * (ch + 1) / NCH;
      double acc = 0.0;
      for (k = k0; k < k1; ++k)
        acc += r[2 * k] * q[2 * k] + r[2 * k + 1] * q[2 * k + 1];
      part[ch] = acc;
    }
    for (ch = 0; ch < NCH; ++ch) al += part[ch];
#pragma omp parallel for schedule(static)
    for (ch = 0; ch < NCH; ++ch) {
      size_t k, k0 = 2 * N * ch / NCH, k1 = 2 * N * (ch + 1) / NCH;
      double acc = 0.0;
      for (k = k0; k < k1; ++k) {
        r[k] -= al * q[k] + bprev * q1[k];
        acc += r[k] * r[k];
        q1[k] = q[k];
      }
      part[ch] = acc;
    }
    for (ch = 0; ch < NCH; ++ch) bb += part[ch];
    b = sqrt(bb);
    alpha[it] = al;
    beta[it] = b;
  }
  free(q);
  free(q1);
  free(r);
}

/* Number of eigenvalues of the tridiagonal below x (Sturm count). */
static int sturm_count(int n, const double *d, const double *e, double x) {
  int i, cnt = 0;
  double q = d[0] - x;
  if (q < 0) ++cnt;
  for (i = 1; i < n; ++i) {
    if (q == 0.0) q = 1e-300;
    q = d[i] - x - e[i - 1] * e[i - 1] / q;
    if (q < 0) ++cnt;
  }
  return cnt;
}

int or_tridiag_eigs(int n, const double *d, const double *e, double lo,
                    double hi, double tol, double *w, int wmax) {
  int klo = sturm_count(n, d, e, lo), khi = sturm_count(n, d, e, hi), k, cnt = 0;
  for (k = klo; k < khi && cnt < wmax; ++k) {
    double a = lo, b = hi;
    while (b - a > tol * (fabs(a) + fabs(b) + 1e-300) && b - a > 1e-300) {
      const double mid = 0.5 * (a + b);
      if (mid <= a || mid >= b) break;
      if (sturm_count(n, d, e, mid) > k) b = mid;
      else a = mid;
    }
    w[cnt++] = 0.5 * (a + b);
  }
  return cnt;
}
