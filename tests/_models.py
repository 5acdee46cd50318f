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
