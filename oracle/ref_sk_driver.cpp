// Harness around the reference's bond routine (h3d_offdiag_diag_corr,
// NEMO_3D/nemo3d/src/ham/mb_ham_spds_micro.c:597-1451), compiled from the
// reference sources where they lie (oracle/ref_build.sh).  Test
// infrastructure only: it produces the strained Slater-Koster block and
// the Boykin diagonal corrections that pin the oracle's or_sk_bond.
//
// stdin, one bond per line:
//   nb anion_row a0 nnv[3] nnv0[3] param[35] strain[33]
// stdout, one line per bond: ho[i][j].r (10x10, row-major) then
//   hd[i][i].r (10): the corrections accumulated from a zero diagonal.
#include <cstdio>

#include "mb_ham_spds_micro.h"

int main() {
  int nb, anion_row;
  double a0;
  while (std::scanf("%d %d %lf", &nb, &anion_row, &a0) == 3) {
    rvectr nnv = Rvectr(3), nnv0 = Rvectr(3);
    rvectr param = Rvectr(35), strain = Rvectr(33);
    for (int k = 0; k < 3; ++k) std::scanf("%lf", &nnv[k]);
    for (int k = 0; k < 3; ++k) std::scanf("%lf", &nnv0[k]);
    for (int k = 0; k < 35; ++k) std::scanf("%lf", &param[k]);
    for (int k = 0; k < 33; ++k) std::scanf("%lf", &strain[k]);
    // basis and spin labels as basis_init sets them (ham3d_final.c:594-616)
    ivectr spin = Ivectr(nb), an = Ivectr(nb), ca = Ivectr(nb);
    for (int i = 0; i < nb; ++i) {
      spin[i] = nb == 20 ? (i < 10 ? 1 : 0) : -1;
      an[i] = i % 10;
      ca[i] = i % 10 + 10;
    }
    cmatrix ho = Cmatrix(nb, nb), hd = Cmatrix(nb, nb);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) {
        ho[i][j].r = ho[i][j].i = 0.0;
        hd[i][j].r = hd[i][j].i = 0.0;
      }
    h3d_offdiag_diag_corr(ho, hd, nnv, nnv0, anion_row ? an : ca,
                          anion_row ? ca : an, spin, param, strain, a0, nb);
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) std::printf("%.17g ", ho[i][j].r);
    for (int i = 0; i < 10; ++i) std::printf("%.17g ", hd[i][i].r);
    std::printf("\n");
    std::fflush(stdout);
  }
  return 0;
}
