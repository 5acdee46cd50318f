"""CPU oracle of the Keating VFF energy (test infrastructure only).

Restates calc_estrain (NEMO_3D/nemo3d/src/base/strain3d.c:1739-1890):
  E = sum_i [ sum_n a_in/(2 d0_in^2) (|r_in|^2 - d0_in^2)^2
            + sum_{n<p} sqrt(b_in b_ip)/(|d0_in||d0_ip|) (r_in.r_ip - d0_in.d0_ip)^2 ]
Pinned by: Keating's elastic constants C11 = (alpha + 3 beta)/a and
C12 = (alpha - beta)/a (Keating 1966, Martin 1970), times the reference's
8/3 prefactor convention, from the energy of a periodic crystal under
hydrostatic, uniaxial and biaxial strain; zero energy and gradient for an
ideal crystal at its own lattice constant; central finite differences
(the GPU gradient is checked against this energy).
"""
import numpy as np


def energy(structure, species, alpha, beta, a0, disp=None):
    s = structure
    u = s.disp if disp is None else disp
    nbr = s.nbr
    A = nbr.shape[0]
    i_idx = np.repeat(np.arange(A), 4).reshape(A, 4)
    ok = nbr >= 0
    j = np.where(ok, nbr, 0)
    q = 0.25 * s.bond_q
    r = s.lattice * q + u[j] - u[i_idx]
    a0p = a0[species[i_idx], species[j]]
    d0 = a0p[..., None] * q
    al = alpha[species[i_idx], species[j]]
    be = beta[species[i_idx], species[j]]
    d0sq = (d0 * d0).sum(-1)
    dl = (r * r).sum(-1) - d0sq
    e = np.where(ok, 0.5 * al / np.where(ok, d0sq, 1.0) * dl * dl, 0.0).sum()
    for n in range(4):
        for p in range(n + 1, 4):
            m = ok[:, n] & ok[:, p]
            K = np.sqrt(be[:, n] * be[:, p]) / np.sqrt(
                np.where(m, d0sq[:, n] * d0sq[:, p], 1.0))
            t = (r[:, n] * r[:, p]).sum(-1) - (d0[:, n] * d0[:, p]).sum(-1)
            e += np.where(m, K * t * t, 0.0).sum()
    return float(e)
