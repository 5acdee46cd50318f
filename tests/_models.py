"""Seeded model builders shared by the tests."""
import numpy as np

from nemo3d_amd import materials, model, structure


def gaas_box(cells=(2, 2, 2), strain_nm=0.0, seed=0, band=materials.BANDS_20_SPIN,
             passivate=True, alloy=False, phi=False):
    """Closed GaAs box, optionally with an InAs/GaAs random cation alloy,
    random displacements (strained bonds) and a random potential."""
    a = 0.56532
    rng = np.random.default_rng(seed)
    elements = ["Ga", "As", "In"]

    def occ(pos, m, cell):
        cat = m < 4
        out = np.where(cat, 0, 1)
        if alloy:
            out = np.where(cat & (rng.random(m.size) < 0.5), 2, out)
        return out

    s = structure.build((1, 1, 1), 1 + np.asarray(cells), occ, elements, (a, a, a))
    if strain_nm:
        s.disp = rng.normal(0.0, strain_nm, size=(s.n_atoms, 3))
    compounds = ["GaAs", "InAs"] if alloy else ["GaAs"]
    ph = rng.normal(0.0, 0.05, size=s.n_atoms) if phi else None
    return model.TBModel(s, compounds, band, passivate=passivate, db_shift=30.0,
                         phi=ph)


def si_box(cells=(3, 3, 3), strain_nm=0.0, seed=0):
    a = 0.543095
    s = structure.binary_box(cells, "Si", "Si", (a, a, a))
    if strain_nm:
        s.disp = np.random.default_rng(seed).normal(0.0, strain_nm, (s.n_atoms, 3))
    return model.TBModel(s, ["Si"], materials.BANDS_10_NOSPIN, passivate=True,
                         db_shift=30.0, source="silicon4.xml")


def random_vectors(n, ncol, seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(ncol, n)) + 1j * rng.normal(size=(ncol, n))


def pq_si48():
    """The structure of the reference's TRACEMIN example Hamiltonian
    (NEMO_3D/TRACEMIN/Examples/pqexample2_full): 3 x 2 x 1 conventional Si
    cells, periodic in all directions (48 atoms, each with four
    neighbours), unstrained, 20-band sp3d5s* + spin-orbit.  The example was
    written with the energies and two-centre integrals of the spinless Si
    record (Bands_10_sp3d5ss_nospin of the decks) and the spin-orbit
    constant of the spin record (0.01989 eV): its onsite energies are the
    spinless record's (-2.14463, 4.2363, 19.12352, 13.79658 eV) and its
    p-p couplings within an atom +-0.01989."""
    a = 0.543095
    s = structure.binary_box((3, 2, 1), "Si", "Si", (a, a, a), periodic=(True, True, True))
    M = model.TBModel(s, ["Si"], materials.BANDS_20_SPIN, source="GaAs_bulk.xml")
    tab = materials.param_init(["Si"], materials.BANDS_10_NOSPIN, source="GaAs_bulk.xml",
                               spinless_view=False)
    spin = materials.find_material("Si", materials.BANDS_20_SPIN, source="GaAs_bulk.xml")
    tab.parmat[:, :, 8] = spin["mb_pv"][8]   # lambda, anion
    tab.parmat[:, :, 9] = spin["mb_pv"][9]   # lambda, cation
    M.tables = tab
    return M


def pq_example_matrix():
    """The example's matrix (dense, from the committed fixture)."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "pqexample2_coo.npz"))
    n = int(d["n"])
    H = np.zeros((n, n), dtype=np.complex128)
    H[d["i"].astype(np.int64), d["j"].astype(np.int64)] = d["re"] + 1j * d["im"]
    return H
