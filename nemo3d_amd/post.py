"""Post-processing of a solve: probability densities of the eigenvectors and
the local strain tensor of the relaxed structure, written as the reference
writes them.

* writePsiSqr (NEMO_3D/nemo3d/src/io/i_3dout_files.c:1425-1520): per atom
  the sum of |psi|^2 over its orbitals, divided by the largest such sum of
  the state; ``nd_wf_<k>`` per atom and ``nd_wf_<k>_unit`` per cell (the
  sum over the cell's atoms), k counted from 1.
* compute_strain_tensor (NEMO_3D/nemo3d/src/ham/mb_ham_spds_micro.c:
  3299-3448): for an atom with all four neighbours, the three edge vectors
  d1 + d2 - d3 - d0, d2 + d3 - d1 - d0, d1 + d3 - d2 - d0 of its bonds
  (bonds labelled by the neighbour's atom index in the cell, the third
  edge negated on the anion sublattice), strained (current lattice and
  displacements) and ideal (each pair's unstrained lattice constant); the
  strain is A A0^-1 - 1, reported as exx, eyy, ezz, exy, exz, eyz.
  Written per atom and per cation as ``strain3d`` / ``strain3d_cation``
  (ham3d_final.c:1173-1223): position x y z and the six components, zeros
  for atoms without four neighbours.
"""
import numpy as np

from . import ndio
from .structure import NNMAP


def psi_squared(psi, nb):
    """Per-atom probability |psi|^2 of one state, divided by its maximum."""
    p = np.abs(np.asarray(psi).reshape(-1, nb)) ** 2
    s = p.sum(axis=1)
    return s / s.max()


def cell_sums(values, structure):
    """Sums of per-atom values over the atoms of each cell, cells in the
    order of their first atom (the Hamiltonian's cell order)."""
    c = np.asarray(structure.cell)
    key = np.ascontiguousarray(c).view([("", c.dtype)] * 3).ravel()
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    out = np.zeros(order.size)
    np.add.at(out, rank[inv], values)
    return out


def write_psi_sqr(prefix, vectors, nb, band_model, structure, wf_indx=0):
    """nd_wf_<k> (per atom) and nd_wf_<k>_unit (per cell) for each state;
    returns the file names."""
    names = []
    for k, v in enumerate(vectors):
        p2 = psi_squared(v, nb)
        name = "%s.nd_wf_%d" % (prefix, wf_indx + k + 1)
        ndio.write(name, "{real psi2}", [(p2.size, "# number of atoms")],
                   ndio._common_user(band_model), p2[:, None])
        cs = cell_sums(p2, structure)
        name_u = "%s.nd_wf_%d_unit" % (prefix, wf_indx + k + 1)
        ndio.write(name_u, "{real psi2_unit}", [(cs.size, "# number of cells")],
                   ndio._common_user(band_model), cs[:, None])
        names += [name, name_u]
    return names


def _block_label_maps():
    """map_mb_to_m (atom, block label) -> the neighbour's atom index in the
    cell, and map_mm_to_n (atom, neighbour atom index) -> slot
    (mb_ham_spds_micro.c:3311-3334)."""
    mb = np.zeros((8, 4), dtype=np.int64)
    for m in range(4):
        mb[m] = [4 + m, 4 + (7 - m) % 4, 4 + (2 + m) % 4, 4 + (5 - m) % 4]
    for m in range(4, 8):
        mb[m] = [(m - 2) % 4, (9 - m) % 4, m % 4, (7 - m) % 4]
    mm = -np.ones((8, 8), dtype=np.int64)
    for m in range(8):
        for n in range(4):
            mm[m, NNMAP[m, n, 3]] = n
    return mb, mm


def strain_tensor(structure, species, a0_pair):
    """(A, 6) strain components exx, eyy, ezz, exy, exz, eyz per atom and
    the (A,) mask of atoms with all four neighbours (the others get zeros).
    a0_pair[s, t]: unstrained lattice constant of the species pair (the
    VFF table, strain.amat)."""
    s = structure
    mb, mm = _block_label_maps()
    A = s.nbr.shape[0]
    ok = (s.nbr >= 0).all(axis=1)
    m = np.asarray(s.m)
    slots = mm[m[:, None], mb[m]]              # (A, 4) slot of block label b
    rows = np.arange(A)[:, None]
    j = s.nbr[rows, slots]
    jj = np.where(j >= 0, j, 0)
    dl = 0.25 * s.bond_q[rows, slots].astype(np.float64)   # units of a
    dR = np.asarray(s.lattice)[None, None, :] * dl + (s.disp[jj] - s.disp[:, None, :])
    alat = a0_pair[np.asarray(species)[:, None], np.asarray(species)[jj]]
    dR0 = alat[..., None] * dl

    def edges(d):
        e = np.stack([d[:, 1] + d[:, 2] - d[:, 3] - d[:, 0],
                      d[:, 2] + d[:, 3] - d[:, 1] - d[:, 0],
                      d[:, 1] + d[:, 3] - d[:, 2] - d[:, 0]], axis=2)  # (A,3,3) cols
        e[~s.is_cation.astype(bool), :, 2] *= -1.0
        return e

    Am, A0 = edges(dR), edges(dR0)
    out = np.zeros((A, 6))
    if ok.any():
        Mt = np.linalg.solve(np.transpose(A0[ok], (0, 2, 1)),
                             np.transpose(Am[ok], (0, 2, 1)))
        M = np.transpose(Mt, (0, 2, 1))          # A A0^-1
        out[ok] = np.stack([M[:, 0, 0] - 1.0, M[:, 1, 1] - 1.0, M[:, 2, 2] - 1.0,
                            M[:, 0, 1], M[:, 0, 2], M[:, 1, 2]], axis=1)
    return out, ok


def write_strain3d(path, path_cation, structure, eps, ok):
    """strain3d / strain3d_cation text files: x y z and the six components
    per atom (per cation), zeros for atoms without four neighbours."""
    p = structure.positions()
    with open(path, "w") as f, open(path_cation, "w") as fc:
        for i in range(p.shape[0]):
            e = eps[i] if ok[i] else None
            line = ("%e %e %e" % tuple(p[i])) + (
                " 0 0 0 0 0 0 \n" if e is None else
                " %e %e %e %e %e %e \n" % tuple(e))
            f.write(line)
            if structure.is_cation[i]:
                fc.write(line)
