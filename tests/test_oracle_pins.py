"""Pins of oracle functions by closed forms that do not come from the code
under test: the 2-D Laplacian spectrum of the reference's Lanczos
benchmark, LAPACK's tridiagonal eigensolver, and the Keating elastic
constants of a homogeneously strained crystal."""
import numpy as np
import pytest
import scipy.linalg

from nemo3d_amd import materials, model, structure, vff
from oracle import oracle, vff_oracle


def laplacian_operator(m, b):
    """The b x b block-tridiagonal 2-D Laplacian of NEMO_3D/eigen/doc/
    lanczos.tex:362-400 (blocks T = tridiag(-1, 4, -1) of size m, -I off
    the diagonal) written in the oracle's row-owner layout: one grid point
    per 'atom', its 10 orbitals uncoupled (the operator is the Laplacian
    times I_10), onsite 4, neighbour blocks -I."""
    A = m * b
    idx = np.arange(A).reshape(b, m)
    nbr = -np.ones((A, 4), dtype=np.int32)
    for r in range(b):
        for c in range(m):
            k = idx[r, c]
            for n, (dr, dc) in enumerate(((0, 1), (0, -1), (1, 0), (-1, 0))):
                rr, cc = r + dr, c + dc
                if 0 <= rr < b and 0 <= cc < m:
                    nbr[k, n] = idx[rr, cc]
    onsite = np.zeros((A, oracle.ONSITE_STRIDE))
    onsite[:, :10] = 4.0
    blocks = np.zeros((A, 4, 10, 10))
    blocks[nbr >= 0] = -np.eye(10)
    return oracle.OracleHamiltonian(10, nbr, onsite, blocks, 0)


def laplacian_spectrum(m, b):
    i = np.arange(1, m + 1)
    j = np.arange(1, b + 1)
    return np.sort((4 - 2 * np.cos(i * np.pi / (m + 1))[:, None]
                    - 2 * np.cos(j * np.pi / (b + 1))[None, :]).ravel())


def test_laplacian_operator_is_the_kat_matrix():
    """The operator above is the lanczos.tex matrix: its dense form, with
    one orbital per point, has the closed-form spectrum
    4 - 2cos(i pi/(m+1)) - 2cos(j pi/(b+1))."""
    m, b = 7, 5
    H = laplacian_operator(m, b)
    D = H.dense()
    w = np.linalg.eigvalsh(D[::10, ::10])
    assert np.abs(w - laplacian_spectrum(m, b)).max() < 1e-13


def test_lanczos_tridiag_finds_laplacian_eigenvalues():
    """or_lanczos_tridiag on the 2-D Laplacian (m = b = 30, the reference's
    benchmark matrix at a small size) from a random start: the Ritz values
    of its tridiagonal in the benchmark's window (4, 8] that have
    converged sit on the closed-form eigenvalues, and each of the five
    largest distinct eigenvalues has a Ritz value on it (a level can be
    momentarily unconverged while a ghost copy forms, so that check does
    not ask for the convergence flag).  A dropped or mis-signed term of
    the three-term recurrence breaks both."""
    m = b = 30
    H = laplacian_operator(m, b)
    exact = laplacian_spectrum(m, b)
    rng = np.random.default_rng(7)
    r0 = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    r0 /= np.linalg.norm(r0)
    k = 400
    a, bt = H.lanczos_tridiag(r0, k)
    # Ritz pairs of T_k; converged when the last component of the Ritz
    # vector times beta_k is tiny (the Lanczos residual bound)
    w, s = scipy.linalg.eigh_tridiagonal(a, bt[:k - 1])
    resid = np.abs(bt[k - 1] * s[-1, :])
    conv = w[(resid < 1e-10) & (w > 4.0)]
    assert conv.size > 10
    dist = np.abs(conv[:, None] - exact[None, :]).min(axis=1)
    assert dist.max() < 1e-9, dist.max()
    top = np.unique(np.round(exact, 9))[-5:]
    for e in top:
        assert np.abs(w - e).min() < 1e-9, e


def test_tridiag_bisection_matches_lapack():
    """or_tridiag_eigs (Sturm bisection) against LAPACK's stebz/stemr via
    scipy on the Laplacian run's tridiagonal and on a random one, in
    several windows (ghost copies included: multiplicities must agree)."""
    m = b = 20
    H = laplacian_operator(m, b)
    rng = np.random.default_rng(3)
    r0 = rng.normal(size=H.n) + 1j * rng.normal(size=H.n)
    r0 /= np.linalg.norm(r0)
    a, bt = H.lanczos_tridiag(r0, 300)
    d2 = rng.normal(size=500)
    e2 = rng.normal(size=500)
    for d, e in ((a, bt[:-1]), (d2, e2[:-1])):
        w_all = scipy.linalg.eigh_tridiagonal(d, e, eigvals_only=True)
        span = w_all.max() - w_all.min()
        # window edges between eigenvalues (an edge on an eigenvalue is
        # decided by rounding)
        mid = 0.5 * (w_all[1:] + w_all[:-1])
        for lo, hi in ((w_all.min() - 1, w_all.max() + 1),
                       (mid[len(mid) // 2], w_all.max() + 1),
                       (mid[len(mid) // 3], mid[len(mid) // 2])):
            w = oracle.tridiag_eigs(d, e, lo, hi)
            ref = w_all[(w_all > lo) & (w_all <= hi)]
            assert w.size == ref.size
            assert np.abs(np.sort(w) - ref).max() < 1e-12 * span


def _periodic_gaas(cells=3):
    a = 0.56532
    s = structure.binary_box((cells,) * 3, "Ga", "As", (a, a, a),
                             periodic=(True, True, True))
    M = model.TBModel(s, ["GaAs"], materials.BANDS_20_SPIN)
    return s, M


@pytest.mark.parametrize("kind", ["hydrostatic", "uniaxial100", "biaxial001"])
def test_vff_energy_gives_keating_elastic_constants(kind):
    """The Keating energy of a periodic GaAs crystal under a small
    homogeneous normal strain equals the elastic energy density with
    Keating's constants C11 = (alpha + 3 beta)/a, C12 = (alpha - beta)/a
    (Keating, Phys. Rev. 145, 637 (1966); Martin, PRB 1, 4005 (1970)),
    times 8/3: the reference drops the model's 3/8 and 3/16 prefactors and
    uses 1/2 on the bond term (NEMO_3D/nemo3d/src/base/strain3d.c:1765-1770,
    1843-1846).  Normal strains of zincblende need no internal relaxation,
    so the homogeneous positions are the equilibrium ones.  Hydrostatic
    strain tests 3 alpha + beta, uniaxial [100] alpha + 3 beta: together
    they pin both prefactors."""
    s, M = _periodic_gaas()
    al, be, a0 = vff.keating_tables(M)
    a = 0.56532
    alpha, beta = al[0, 1], be[0, 1]
    C11 = (alpha + 3 * beta) / a
    C12 = (alpha - beta) / a
    eps = {"hydrostatic": (1, 1, 1), "uniaxial100": (1, 0, 0),
           "biaxial001": (1, 1, 0)}[kind]
    V = s.n_atoms / 8 * a ** 3
    for e0 in (2e-4, 1e-4):
        ev = np.asarray(eps) * e0
        s.lattice = a * (1 + ev)
        E = vff_oracle.energy(s, M.species, al, be, a0)
        W = 0.5 * C11 * (ev ** 2).sum() + C12 * (ev[0] * ev[1] + ev[1] * ev[2]
                                                 + ev[0] * ev[2])
        # finite-strain terms are O(eps) relative
        assert abs(E / V - 8.0 / 3.0 * W) <= 5 * e0 * (8.0 / 3.0 * W), (kind, E / V, W)


def test_strained_keating_reduces_to_keating():
    """VFF_keating_strained with zero anharmonicity constants (A = B = C =
    0: alpha_eff = alpha, beta_eff = the first leg's beta) is the harmonic
    Keating energy on a binary crystal (both legs' beta equal), at random
    displacements; with the deck's constants it differs."""
    a = 0.56532
    s = structure.binary_box((3, 3, 3), "Ga", "As", (a, a, a))
    M = model.TBModel(s, ["GaAs"], materials.BANDS_20_SPIN)
    s.disp = np.random.default_rng(4).normal(0, 0.01, (s.n_atoms, 3))
    al, be, a0 = vff.keating_tables(M)
    aa, ba, ca = vff.anharmonic_tables(M)
    z = np.zeros_like(aa)
    e_h = vff_oracle.energy(s, M.species, al, be, a0)
    e_0 = vff_oracle.energy(s, M.species, al, be, a0, model="VFF_keating_strained",
                            aa=z, ba=z + 1e-300, ca=z + 1e-300)
    e_s = vff_oracle.energy(s, M.species, al, be, a0, model="VFF_keating_strained",
                            aa=aa, ba=ba, ca=ca)
    assert abs(e_0 - e_h) <= 1e-12 * e_h
    assert abs(e_s - e_h) > 1e-3 * e_h
