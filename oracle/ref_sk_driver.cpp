// Harness around two routines of the reference's sp3d5s* Hamiltonian,
// compiled from the reference sources where they lie (oracle/ref_build.sh).
// Test infrastructure only: its outputs pin the oracle's or_sk_bond and
// or_onsite.
//
// Bond mode (no argument): the strained Slater-Koster block and the Boykin
// diagonal corrections of h3d_offdiag_diag_corr
// (NEMO_3D/nemo3d/src/ham/mb_ham_spds_micro.c:597-1451).
//   stdin, one bond per line:
//     nb anion_row a0 nnv[3] nnv0[3] param[35] strain[33]
//   stdout, one line per bond: ho[i][j].r (10x10, row-major) then
//     hd[i][i].r (10): the corrections accumulated from a zero diagonal.
//
// Onsite mode (argument "diag"): the onsite block of h3d_diag
// (mb_ham_spds_micro.c:1453-1505: diagonal energies and spin-orbit).
//   stdin, one atom per line:  nb anion param[35]
//   stdout, one line per atom: h[i][j].r h[i][j].i for i, j < nb
//     (row-major, interleaved).
#include <cstdio>
#include <cstring>

#include "mb_ham_spds_micro.h"

// basis and spin labels as basis_init sets them (ham3d_final.c:594-616):
// anion orbitals 0-9, cation orbitals 10-19, spin up rows 0-9
static void labels(int nb, ivectr spin, ivectr an, ivectr ca) {
  for (int i = 0; i < nb; ++i) {
    spin[i] = nb == 20 ? (i < 10 ? 1 : 0) : -1;
    an[i] = i % 10;
    ca[i] = i % 10 + 10;
  }
}

static int run_diag() {
  int nb, anion;
  while (std::scanf("%d %d", &nb, &anion) == 2) {
    rvectr param = Rvectr(35);
    for (int k = 0; k < 35; ++k) std::scanf("%lf", &param[k]);
    ivectr spin = Ivectr(nb), an = Ivectr(nb), ca = Ivectr(nb);
    labels(nb, spin, an, ca);
    cmatrix h = Cmatrix(nb, nb);
    h3d_diag(h, anion ? an : ca, spin, param, nb);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j)
        std::printf("%.17g %.17g ", h[i][j].r, h[i][j].i);
    std::printf("\n");
    std::fflush(stdout);
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc > 1 && std::strcmp(argv[1], "diag") == 0) return run_diag();
  int nb, anion_row;
  double a0;
  while (std::scanf("%d %d %lf", &nb, &anion_row, &a0) == 3) {
    rvectr nnv = Rvectr(3), nnv0 = Rvectr(3);
    rvectr param = Rvectr(35), strain = Rvectr(33);
    for (int k = 0; k < 3; ++k) std::scanf("%lf", &nnv[k]);
    for (int k = 0; k < 3; ++k) std::scanf("%lf", &nnv0[k]);
    for (int k = 0; k < 35; ++k) std::scanf("%lf", &param[k]);
    for (int k = 0; k < 33; ++k) std::scanf("%lf", &strain[k]);
    ivectr spin = Ivectr(nb), an = Ivectr(nb), ca = Ivectr(nb);
    labels(nb, spin, an, ca);
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
