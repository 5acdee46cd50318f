"""CPU oracle of the Keating VFF energy (test infrastructure only).

Restates calc_estrain (NEMO_3D/nemo3d/src/base/strain3d.c:1739-1890):
  E = sum_i [ sum_n a_in/(2 d0_in^2) (|r_in|^2 - d0_in^2)^2
            + sum_{n<p} b_inp/(|d0_in||d0_ip|) (r_in.r_ip - d0_in.d0_ip)^2 ]
with b_inp = sqrt(b_in b_ip) for StrainModel VFF_keating.  For
VFF_keating_strained a_in and b_inp are StrainParam::alpha_eff /
beta_eff (StrainParam.c:60-214, acoustic constants aa, ba, ca, truncation
constant 0.001): a (1 - A (|r|^2/d0^2 - 1)) per bond and
b_in (1 - B (cos + 1/3)) (1 - C (rd - 1)) per angle (b_in of the leg with
the lower slot, B and C geometric means over the two legs), each truncated
near its extremum.  ``frozen`` evaluates those constants at other
displacements: the reference's gradient is the derivative of E with the
constants frozen at the current point.

Pinned by: Keating's elastic constants C11 = (alpha + 3 beta)/a and
C12 = (alpha - beta)/a (Keating 1966, Martin 1970), times the reference's
8/3 prefactor convention, from the energy of a periodic crystal under
hydrostatic, uniaxial and biaxial strain; zero energy and gradient for an
ideal crystal at its own lattice constant; central finite differences
(the GPU gradient is checked against this energy with each bond term once,
frozen constants for the strained model); the strained model with
A = B = C = 0 is the harmonic one; the reference's InGaAs_QD test run
(its relaxed structure enters both printed eigenvalues).  The scaling laws of the strained model themselves are
parity unpinned apart from those reductions (no reference output holds
them in isolation).
"""
import numpy as np

NEAR_MAX = 0.001


def alpha_eff(al0, A, r2, d2):
    # absent bonds (d2 = 0) give non-finite values here that the callers mask
    delt = 1.0 - NEAR_MAX
    with np.errstate(divide="ignore", invalid="ignore"):
        inside = r2 < d2 * (1.0 + delt * 2.0 / (3.0 * A))
        return np.where(inside, al0 * (1.0 - A * (r2 / d2 - 1.0)),
                        al0 * (1.0 - delt / 1.5))


def beta_eff(be0, B, C, d1, d2, r1, r2):
    # absent legs give non-finite values that the callers mask
    with np.errstate(divide="ignore", invalid="ignore"):
        return _beta_eff(be0, B, C, d1, d2, r1, r2)


def _beta_eff(be0, B, C, d1, d2, r1, r2):
    d10 = np.sqrt((d1 * d1).sum(-1))
    d20 = np.sqrt((d2 * d2).sum(-1))
    r10 = np.sqrt((r1 * r1).sum(-1))
    r20 = np.sqrt((r2 * r2).sum(-1))
    rd = r10 * r20 / (d10 * d20)
    tt = (r1 * r2).sum(-1) / (r10 * r20)
    delt, delh = 1.0 + NEAR_MAX, 1.0 - NEAR_MAX
    with np.errstate(divide="ignore", invalid="ignore"):
        ang = tt > delt / 3.0 * (2.0 * (1.0 / B - 1.0 / 3.0) - 1.0 / (3.0 * rd))
        lng = rd > 1.0 + delh * 2.0 / (3.0 * C)
    b_ang = be0 * ((1.0 - 2.0 * delt / 3.0) +
                   B / 3.0 * (delt / 3.0 * (1.0 / rd + 2.0) - 1.0)) * (1.0 - C * (rd - 1.0))
    b_len = be0 * (1.0 - B * (tt + 1.0 / 3.0)) * (1.0 - delh * 2.0 / 3.0)
    b_std = be0 * (1.0 - B * (tt + 1.0 / 3.0)) * (1.0 - C * (rd - 1.0))
    out = np.where(ang, b_ang, b_std)
    return np.where(lng, np.where(ang, np.maximum(b_ang, b_len), b_len), out)


def _bonds(s, u):
    nbr = s.nbr
    A = nbr.shape[0]
    i_idx = np.repeat(np.arange(A), 4).reshape(A, 4)
    ok = nbr >= 0
    j = np.where(ok, nbr, 0)
    q = 0.25 * s.bond_q
    r = s.lattice * q + u[j] - u[i_idx]
    return i_idx, j, ok, q, r


def constants(structure, species, alpha, beta, a0, disp=None, model="VFF_keating",
              aa=None, ba=None, ca=None):
    """Per-bond alpha [A,4] and per-angle beta [A,4,4] (n < p) at disp."""
    s = structure
    u = s.disp if disp is None else disp
    i_idx, j, ok, q, r = _bonds(s, u)
    ti, tj = species[i_idx], species[j]
    d0 = a0[ti, tj][..., None] * q
    al = alpha[ti, tj]
    if model == "VFF_keating_strained":
        al = alpha_eff(al, aa[ti, tj], (r * r).sum(-1), (d0 * d0).sum(-1))
    A = s.nbr.shape[0]
    bet = np.zeros((A, 4, 4))
    for n in range(4):
        for p in range(n + 1, 4):
            b1, b2 = beta[ti[:, n], tj[:, n]], beta[ti[:, p], tj[:, p]]
            if model == "VFF_keating_strained":
                B = np.sqrt(ba[ti[:, n], tj[:, n]] * ba[ti[:, p], tj[:, p]])
                C = np.sqrt(ca[ti[:, n], tj[:, n]] * ca[ti[:, p], tj[:, p]])
                bet[:, n, p] = beta_eff(np.sqrt(b1 * b1), B, C, d0[:, n], d0[:, p],
                                        r[:, n], r[:, p])
            else:
                bet[:, n, p] = np.sqrt(b1 * b2)
    return al, bet


def energy(structure, species, alpha, beta, a0, disp=None, model="VFF_keating",
           aa=None, ba=None, ca=None, frozen=None, bond_copies=2):
    """E at disp; frozen: displacements at which the model's constants are
    evaluated (default disp itself).  bond_copies=2: each bond term from
    both of its atoms (the energy calc_estrain reports); 1: each bond once,
    the function whose derivative calc_estrain's gradient is (an atom gets
    the derivative of its own copy of a bond term only, strain3d.c:
    1806-1830)."""
    s = structure
    u = s.disp if disp is None else disp
    al, bet = constants(s, species, alpha, beta, a0, u if frozen is None else frozen,
                        model, aa, ba, ca)
    i_idx, j, ok, q, r = _bonds(s, u)
    d0 = a0[species[i_idx], species[j]][..., None] * q
    d0sq = (d0 * d0).sum(-1)
    dl = (r * r).sum(-1) - d0sq
    e = 0.5 * bond_copies * np.where(ok, 0.5 * al / np.where(ok, d0sq, 1.0) * dl * dl,
                                     0.0).sum()
    for n in range(4):
        for p in range(n + 1, 4):
            m = ok[:, n] & ok[:, p]
            K = bet[:, n, p] / np.sqrt(np.where(m, d0sq[:, n] * d0sq[:, p], 1.0))
            t = (r[:, n] * r[:, p]).sum(-1) - (d0[:, n] * d0[:, p]).sum(-1)
            e += np.where(m, K * t * t, 0.0).sum()
    return float(e)
