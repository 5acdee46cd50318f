"""The CPU oracle against the reference's golden data and closed forms."""
import os
import sys

import numpy as np
import pytest

from nemo3d_amd import materials, model, structure
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def gaas_bulk_model():
    a = 0.56532
    s = structure.binary_box((1, 1, 1), "Ga", "As", (a, a, a),
                             periodic=(True, True, True))
    return s, model.TBModel(s, ["GaAs"], materials.BANDS_20_SPIN,
                            source="GaAs_bulk.xml")


def test_gaas_bulk_golden_spectrum():
    """Periodic GaAs, 20-band sp3d5s* + SO: 160 eigenvalues at 40 k-points
    (nemo3d/example/TestRun/GaAs_bulk.old_nd_Ek_ascii, 13 digits)."""
    g = np.load(os.path.join(GOLDEN, "gaas_bulk_ek.npz"))
    s, M = gaas_bulk_model()
    H = oracle.assemble(M.arrays())
    assert H.nsurf == 0
    worst = 0.0
    for kx in np.unique(g["kx"]):
        ph = np.exp(1j * np.pi * kx * s.wrapped[:, :, 0])
        Hd = H.dense(phase=ph)
        assert np.abs(Hd - Hd.conj().T).max() < 1e-14
        w = np.linalg.eigvalsh(Hd)
        ref = np.sort(g["E"][g["kx"] == kx])
        worst = max(worst, np.abs(w - ref).max())
    # the golden prints 13 significant digits of values up to ~40 eV
    assert worst < 2e-11, worst


def test_gaas_bulk_golden_detects_sign_error():
    """A flipped sign of one two-centre integral must break the golden."""
    g = np.load(os.path.join(GOLDEN, "gaas_bulk_ek.npz"))
    s, M = gaas_bulk_model()
    arr = M.arrays()
    arr["parmat"] = arr["parmat"].copy()
    arr["parmat"][:, :, 24] *= -1.0  # p_a d_c sigma
    H = oracle.assemble(arr)
    kx = np.unique(g["kx"])[5]
    ph = np.exp(1j * np.pi * kx * s.wrapped[:, :, 0])
    w = np.linalg.eigvalsh(H.dense(phase=ph))
    ref = np.sort(g["E"][g["kx"] == kx])
    assert np.abs(w - ref).max() > 1e-3


def test_silicon4_structure_matches_reference():
    """64,000-atom closed Si box: positions and the neighbour table equal
    the reference's nd_rAtom / nd_nbrIndx output."""
    f = np.load(os.path.join(GOLDEN, "silicon4_meta.npz"))
    a = 0.543095
    s = structure.binary_box((20, 20, 20), "Si", "Si", (a, a, a))
    assert s.n_atoms == 64000
    assert np.array_equal(s.nbr, f["nbr"])
    p = s.positions()
    assert np.array_equal(p[:64], f["r_first"])
    assert np.array_equal(p[-64:], f["r_last"])
    assert np.allclose(p.sum(axis=0), f["r_sum"], rtol=1e-14)
    assert int((s.nbr < 0).sum()) == int(f["n_dangling"])


def _nd(path):
    b = open(path, "rb").read()
    return b[b.find(b"<\\HDR>") + 6:]


def test_silicon4_rayleigh_quotient(reference_root):
    """Shipped eigenvector 3 of silicon4 (the one the reference left
    normalised) is an eigenvector of the oracle H: Rayleigh quotient equals
    nd_Ek[2] and the residual is small.  Uses the potential file SQD_V and
    30 eV dangling-bond passivation as the run did."""
    t = os.path.join(reference_root, "PostNEMO3D/bin/test/")
    a = 0.543095
    s = structure.binary_box((20, 20, 20), "Si", "Si", (a, a, a))
    M = model.TBModel(s, ["Si"], materials.BANDS_10_NOSPIN, passivate=True,
                      db_shift=30.0, source="silicon4.xml",
                      phi=np.loadtxt(t + "SQD_V"))
    H = oracle.assemble(M.arrays())
    assert H.nsurf == 4681  # sum of Natom_surface over ranks in output.out
    E = np.frombuffer(_nd(t + "silicon4.nd_Ek"), "<f8").reshape(-1, 4)[:, 3]
    ev = np.frombuffer(_nd(t + "silicon4.nd_evec_3"), "<f8").reshape(-1, 2)
    psi = ev[:, 0] + 1j * ev[:, 1]
    hp = H.hv(psi)
    rq = np.vdot(psi, hp).real / np.vdot(psi, psi).real
    res = np.linalg.norm(hp - rq * psi) / np.linalg.norm(psi)
    assert abs(rq - E[2]) < 3e-7, (rq, E[2])
    assert res < 3e-6
    # without the potential the same vector is far from an eigenvector
    M0 = model.TBModel(s, ["Si"], materials.BANDS_10_NOSPIN, passivate=True,
                       db_shift=30.0, source="silicon4.xml")
    rq0 = np.vdot(psi, oracle.assemble(M0.arrays()).hv(psi)).real / \
        np.vdot(psi, psi).real
    assert abs(rq0 - E[2]) > 1e-3


def test_ideal_bond_has_no_strain_terms():
    """Ideal bond length and direction: scaling factor 1, Boykin terms 0.

    Up to the reference's truncated sqrt(3) literal (1.73205080757), which
    makes d0/d = 1 - 7e-13 for an ideal bond and leaves ~1e-11 eV terms."""
    tab = materials.param_init(["GaAs"], materials.BANDS_20_SPIN)
    par, strn = tab.parmat[0, 1], tab.strnmat[0, 1]
    a0 = par[31]
    nnv0 = 0.25 * a0 * np.array([1.0, 1.0, 1.0])
    T, ca, cc = oracle.sk_bond(par, strn, a0, nnv0, nnv0)
    assert np.abs(ca).max() < 1e-10 and np.abs(cc).max() < 1e-10
    T1, _, _ = oracle.sk_bond(par, None, a0, nnv0, nnv0)
    assert np.allclose(T, T1, rtol=1e-11, atol=0)


def test_bond_length_power_law():
    """Stretching a bond along its ideal direction scales each two-centre
    integral by (d0/d)^eta: the s-s entry is V_sss * (d0/d)^eta_sss and the
    s*-s* entry V_ststs * (d0/d)^eta_ststs."""
    tab = materials.param_init(["InAs"], materials.BANDS_20_SPIN)
    par, strn = tab.parmat[0, 1], tab.strnmat[0, 1]
    a0 = par[31]
    u = np.array([1.0, -1.0, -1.0]) / np.sqrt(3.0)
    d0 = 0.25 * 1.73205080757 * a0
    for f in (0.93, 1.0, 1.07):
        T, ca, cc = oracle.sk_bond(par, strn, a0, f * d0 * u, 0.25 * a0 * u * np.sqrt(3.0))
        brat = 1.0 / f
        assert T[1, 1] == pytest.approx(par[10] * brat ** strn[2], rel=1e-12)
        assert T[0, 0] == pytest.approx(par[11] * brat ** strn[1], rel=1e-12)
        # p-p along the bond: l^2 pps + (1-l^2) ppp with l = 1/sqrt(3)
        pp = (par[22] * brat ** strn[4] + 2 * par[23] * brat ** strn[5]) / 3
        assert T[2, 2] == pytest.approx(pp, rel=1e-12)
        if f != 1.0:  # strained: Boykin correction of the s row is
            # sum over cation orbitals ((V0^2 - V^2) c / (Ea+Ec)) c_pair
            assert abs(ca[1]) > 0


def test_closed_box_hermitian_and_surface():
    """Small closed GaAs box: assembled H is Hermitian, passivation changes
    only surface atoms' onsite blocks."""
    a = 0.56532
    s = structure.binary_box((2, 2, 2), "Ga", "As", (a, a, a))
    M = model.TBModel(s, ["GaAs"], materials.BANDS_20_SPIN, passivate=True,
                      db_shift=30.0)
    H = oracle.assemble(M.arrays())
    Hd = H.dense()
    assert np.abs(Hd - Hd.conj().T).max() < 1e-13
    assert H.nsurf == int(s.surface_mask().sum())
    M0 = model.TBModel(s, ["GaAs"], materials.BANDS_20_SPIN, passivate=False)
    H0 = oracle.assemble(M0.arrays())
    changed = np.abs(H.onsite - H0.onsite).max(axis=1) > 0
    assert np.array_equal(changed, s.surface_mask())


def _ref_sk_check(data):
    """Oracle bond (or_sk_bond) against the reference routine's outputs:
    the anion-row block is T^T in the reference's storage, the cation-row
    block T; the diagonal corrections go to the row atom."""
    worst_t = worst_c = 0.0
    for k in range(len(data["nb"])):
        T, ca, cc = oracle.sk_bond(data["par"][k], data["strain"][k], data["a0"][k],
                                   data["nnv"][k], data["nnv0"][k])
        want = T.T if data["anion_row"][k] else T
        worst_t = max(worst_t, np.abs(data["ho"][k] - want).max())
        corr = ca if data["anion_row"][k] else cc
        worst_c = max(worst_c, np.abs(data["hd"][k] - corr).max())
    return worst_t, worst_c


def test_strained_bonds_match_reference_fixture():
    """Strained SK blocks + Boykin corrections of random bonds, as written by
    the reference's own h3d_offdiag_diag_corr (tests/golden/
    make_ref_sk_fixture.py), both band models and row roles."""
    d = np.load(os.path.join(GOLDEN, "ref_sk_bonds.npz"))
    assert np.abs(d["hd"]).max() > 1e-2  # the strain corrections are exercised
    worst_t, worst_c = _ref_sk_check(d)
    assert worst_t < 1e-13 and worst_c < 1e-14, (worst_t, worst_c)


def test_strained_bonds_detect_planted_error():
    """A mistake planted in the oracle's input (one strain exponent) breaks
    the agreement with the reference outputs."""
    d = dict(np.load(os.path.join(GOLDEN, "ref_sk_bonds.npz")))
    d["strain"] = d["strain"].copy()
    d["strain"][:, 4] += 0.1  # eta of pp sigma
    worst_t, worst_c = _ref_sk_check(d)
    assert worst_t > 1e-4


def test_strained_bonds_match_reference_harness(tmp_path):
    """Fresh random bonds through the compiled reference routine
    (oracle/_ref/ref_sk, built by oracle/ref_build.sh where the reference
    tree is present)."""
    harness = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "ref_sk")
    if not os.path.exists(harness):
        pytest.skip("reference harness not built (no reference tree)")
    sys.path.insert(0, GOLDEN)
    import make_ref_sk_fixture as mk
    cs = mk.cases(seed=11, n=16)
    ho, hd = mk.run(cs, harness)
    d = dict(nb=np.array([c[0] for c in cs]), anion_row=np.array([c[1] for c in cs]),
             a0=np.array([c[2] for c in cs]), nnv=np.array([c[3] for c in cs]),
             nnv0=np.array([c[4] for c in cs]), par=np.array([c[5] for c in cs]),
             strain=np.array([c[6] for c in cs]), ho=ho, hd=hd)
    worst_t, worst_c = _ref_sk_check(d)
    assert worst_t < 1e-13 and worst_c < 1e-14, (worst_t, worst_c)


def _ref_onsite_worst(data):
    """Oracle onsite (or_onsite, applied through or_hv) against the
    reference h3d_diag blocks of the fixture."""
    worst = 0.0
    for k in range(len(data["nb"])):
        nb = int(data["nb"][k])
        d, lam = oracle.onsite(data["par"][k], data["anion"][k])
        blk = oracle.onsite_block(nb, d, lam)
        worst = max(worst, np.abs(blk - data["h"][k][:nb, :nb]).max())
    return worst


def test_onsite_matches_reference_fixture():
    """Onsite blocks (diagonal energies and the spin-orbit entries) of the
    GaAs/InAs records, their mean and random parameter vectors, as written
    by the reference's own h3d_diag (mb_ham_spds_micro.c:1453-1505), both
    band models and both roles."""
    d = np.load(os.path.join(GOLDEN, "ref_onsite.npz"))
    assert np.abs(d["h"].imag).max() > 1e-3  # spin-orbit exercised
    assert _ref_onsite_worst(d) < 1e-14


def test_onsite_fixture_detects_planted_error():
    """Swapping two parameter entries in the oracle's input (the s and s*
    energies of the anion) breaks the agreement."""
    d = dict(np.load(os.path.join(GOLDEN, "ref_onsite.npz")))
    d["par"] = d["par"].copy()
    d["par"][:, [0, 1]] = d["par"][:, [1, 0]]
    assert _ref_onsite_worst(d) > 1e-3


def test_onsite_matches_reference_harness():
    """Fresh random parameter vectors through the compiled h3d_diag."""
    harness = os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref", "ref_sk")
    if not os.path.exists(harness):
        pytest.skip("reference harness not built (no reference tree)")
    sys.path.insert(0, GOLDEN)
    import make_ref_sk_fixture as mk
    cs = mk.onsite_cases(seed=12)
    h = mk.run_onsite(cs, harness)
    d = dict(nb=np.array([c[0] for c in cs]), anion=np.array([c[1] for c in cs]),
             par=np.array([c[2] for c in cs]), h=h)
    assert _ref_onsite_worst(d) < 1e-14


@pytest.mark.parametrize("band", ["spin", "nospin"])
def test_mixed_species_onsite_is_the_neighbour_mean(band):
    """Onsite terms of an anion with cations of two species around it.
    H_Diag_Wrapper (h_cvectr_mult.c:854-918) averages parmat[c][a] and the
    strain shift over the present neighbours and h3d_diag is linear in the
    parameters (pinned above against the reference routine), so the
    onsite block of an As atom with k In and (n - k) Ga neighbours is the
    k/n : (n-k)/n mix of its blocks in pure InAs and pure GaAs, also at a
    (001) surface with n = 2.  Taking one neighbour's record, or dividing
    by 4 at a surface, breaks this."""
    bm = materials.BANDS_20_SPIN if band == "spin" else materials.BANDS_10_NOSPIN
    a = 0.56532
    s = structure.build((0, 0, 0), (3, 3, 3), lambda pos, m, cell: np.where(
        m < 4, 1, 2), ["In", "Ga", "As"], (a, a, a))
    rng = np.random.default_rng(5)
    s.species = s.species.copy()
    cat = np.nonzero(s.is_cation)[0]
    s.species[cat] = rng.choice([0, 1], cat.size)  # In or Ga per cation site

    def arrays(species_of_cations=None):
        s2 = structure.Structure(s.elements, s.species.copy(), s.is_cation, s.nbr,
                                 s.bond_q, s.disp, s.cell, s.m, s.lattice,
                                 s.periodic, s.wrapped)
        if species_of_cations is not None:
            s2.species[cat] = species_of_cations
        return model.TBModel(s2, ["InAs", "GaAs"], bm).arrays()

    H = oracle.assemble(arrays())
    Hin = oracle.assemble(arrays(0))
    Hga = oracle.assemble(arrays(1))
    checked = set()
    for i in np.nonzero(~s.is_cation.astype(bool))[0]:
        nb_i = s.nbr[i][s.nbr[i] >= 0]
        n = nb_i.size
        k = int((s.species[nb_i] == 0).sum())
        if n < 2:
            continue
        want = (k * Hin.onsite[i] + (n - k) * Hga.onsite[i]) / n
        # surface s-p couplings and passivation do not depend on species
        assert np.abs(H.onsite[i] - want).max() < 1e-12, (i, n, k)
        checked.add((n, k))
    # interior and surface anions with mixed neighbourhoods were seen
    assert {(4, 1), (4, 2), (4, 3)} <= checked
    assert any(n < 4 and 0 < k < n for n, k in checked)


def test_tracemin_example_hamiltonian():
    """The reference's TRACEMIN example Hamiltonian (TRACEMIN/Examples/
    pqexample2_full, a 960 x 960 matrix NEMO-3D wrote for the example:
    48-atom periodic Si, 20-band + spin-orbit, printed to 13 significant
    digits): the oracle's assembly of the same structure equals it entry
    for entry to the print precision, SK couplings, spin-orbit pairs,
    periodic neighbours and row order included."""
    import _models
    M = _models.pq_si48()
    Hd = oracle.assemble(M.arrays()).dense()
    Hc = _models.pq_example_matrix()
    assert Hd.shape == Hc.shape == (960, 960)
    assert np.abs(Hd - Hc).max() <= 5e-12
    assert np.count_nonzero(np.abs(Hc) > 0) == 34560
